// One D2Q9 time step on the GPU, and the vl_sim C ABI.
//
// A step (reference: Sim::step, src/sim.rs:391-420) is ONE launch of k_step:
//   * boundary workgroups (lowest ids, dispatched first): the outlet-controller density
//     offset from A (src/sim.rs:443-455), then one boundary node per thread -- pull, link
//     rules, collide, A -> B, and each force link's momentum-exchange term
//     (src/sim.rs:456-507); workgroup 0 also reduces the previous step's force terms with a
//     fixed-order FP64 tree (double-buffered, so no cross-workgroup synchronisation);
//   * bulk workgroups: pull-stream + collide of every interior fluid node, A -> B
//     (src/sim.rs:700-735 with src/lattice.rs:80-156);
// then A and B swap.  Both halves write disjoint nodes of B and read only A, exactly the
// independence the reference's pull scheme guarantees (src/sim.rs:5-9).
//
// Storage classes: F64 keeps f_i and uses the reference's operation order, so with
// -ffp-contract=off the populations are bit-identical to the reference.  F32 / F16 keep the
// deviation d_i = f_i - w_i (F16 scaled by 2^10) and collide in FP32.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "lbm_math.h"
#include "sim_internal.h"
#include "step_device.h"

namespace vl {

// Small domains: the whole step in one launch (boundary workgroups first).
template <class S, int M>
__global__ __launch_bounds__(STEP_T) void k_step(const typename S::T* __restrict__ A, typename S::T* __restrict__ B,
                                                 const uint8_t* __restrict__ flag, const int2* __restrict__ rows,
                                                 Geom g, Coll c, DevTables t, StepArgs sa) {
  __shared__ double red[STEP_T];
  const int blk = (int)blockIdx.x;
  if (blk < sa.nlb)
    boundary_block<S, M>(blk, A, B, g, c, t, sa, red);
  else
    bulk_node<S, M>(blk - sa.nlb, STEP_T, sa.xblocks, A, B, flag, rows, g, c);
}

// Large domains: the boundary part and the bulk as two kernels on two streams, so each gets
// its own register budget (the bulk stays lean enough for 8 waves per SIMD) and the short
// boundary kernel runs concurrently with the bulk.
template <class S, int M>
__global__ __launch_bounds__(STEP_T) void k_boundary(const typename S::T* __restrict__ A,
                                                     typename S::T* __restrict__ B, Geom g, Coll c, DevTables t,
                                                     StepArgs sa) {
  __shared__ double red[STEP_T];
  boundary_block<S, M>((int)blockIdx.x, A, B, g, c, t, sa, red);
}

constexpr int BULK_T = 512;

// V = 0: one node per thread with the x-shifted populations staged through LDS (bulk_lds).
template <class S, int M, int V>
__global__ __launch_bounds__(BULK_T) void k_bulk(const typename S::T* __restrict__ A, typename S::T* __restrict__ B,
                                                 const uint8_t* __restrict__ flag, const int2* __restrict__ rows,
                                                 Geom g, Coll c, int xblocks) {
  if constexpr (V == 0)
    bulk_lds<S, M, BULK_T>((int)blockIdx.x, xblocks, A, B, flag, rows, g, c);
  else if constexpr (V == 1)
    bulk_node<S, M>((int)blockIdx.x, BULK_T, xblocks, A, B, flag, rows, g, c);
  else
    bulk_vec<S, M, V>((int)blockIdx.x, BULK_T, xblocks, A, B, flag, rows, g, c);
}

// K steps of the covered interior in one pass: one task per wave, W waves per workgroup, V
// columns per lane, D rows of loads in flight; JOIN: joined tasks only (a workgroup's waves
// sweep one strip), else single-wave tasks only (the two forms are separate kernels, each with
// its own register allocation).  FP32 and FP16 storage keep 4 waves per SIMD (at most 128
// registers; the FP16 pass needed 130 unbounded, and 3 waves ran 11 % slower); FP64 needs about
// 185 and runs at 2.
template <class S, int M, int W, int V, int D, int K, bool JOIN>
__global__ __launch_bounds__(W * TB_LANES, sizeof(typename S::T) < 8 ? 4 : 1) void k_pass(const typename S::T* __restrict__ A,
                                                      PassOut<typename S::T> po, const int4* __restrict__ tasks,
                                                      int ntasks, Geom g, Coll c) {
  const int w = __builtin_amdgcn_readfirstlane((int)blockIdx.x * W + (int)(threadIdx.x / TB_LANES));
  if constexpr (JOIN) {
    __shared__ typename S::C xch[2 * W * TB_XCH];  // edge values of the waves
    if (w >= ntasks) return;  // never within a workgroup: joined tasks fill whole workgroups
    multi_step_strip<S, M, V, D, K, true>(tasks[w], A, po, g, c, xch,
                                          __builtin_amdgcn_readfirstlane((int)(threadIdx.x / TB_LANES)), W);
  } else {
    if (w >= ntasks) return;
    multi_step_strip<S, M, V, D, K, false>(tasks[w], A, po, g, c);
  }
}

// One single step of the nodes a K-step pass does not cover, in one launch: workgroups
// [0, nlb) are the boundary part of the step (as in k_boundary), the rest take the listed
// interior nodes.
template <class S, int M>
__global__ __launch_bounds__(STEP_T) void k_phase(const typename S::T* __restrict__ A, typename S::T* __restrict__ B,
                                                  const uint32_t* __restrict__ nodes, int n, Geom g, Coll c,
                                                  DevTables t, StepArgs sa) {
  __shared__ double red[STEP_T];
  const int blk = (int)blockIdx.x;
  if (blk < sa.nlb)
    boundary_block<S, M>(blk, A, B, g, c, t, sa, red);
  else
    listed_node<S, M>((blk - sa.nlb) * STEP_T + (int)threadIdx.x, n, nodes, A, B, g, c);
}

// Advances the window counter at the end of a replayed window of steps.
__global__ void k_advance(int64_t* kbase, int64_t by) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *kbase += by;
}

// Group forces of the last step of a run (the in-step reduction covers the ones before).
__global__ __launch_bounds__(STEP_T) void k_force(const double* __restrict__ fc, const int32_t* __restrict__ gfirst,
                                                  int ngroups, double* __restrict__ out) {
  __shared__ double red[STEP_T];
  reduce_forces(fc, gfirst, ngroups, out, red);
}

// Plain device copy (one 16-byte element per thread, the fastest simple copy form measured on
// MI355X), the bandwidth yardstick reported beside the kernel's own figure.
__global__ __launch_bounds__(256) void k_copy(const float4* __restrict__ src, float4* __restrict__ dst, size_t n) {
  const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (i < n) dst[i] = src[i];
}

// ------------------------------------------------------------------------------------------
// Host side
// ------------------------------------------------------------------------------------------
#define VL_TRY(x)                 \
  do {                            \
    vl_status st_ = (x);          \
    if (st_ != VL_OK) return st_; \
  } while (0)

static const double SQRT_CS2 = std::sqrt(1.0 / 3.0);
// Steps per run chunk: controls go up and forces come back once per chunk.
static constexpr int64_t RUN_CHUNK = 4096;

// Step arguments of the step at time `t`, step `k_local` of a window whose first chunk index is
// *kbase.
static StepArgs step_args(vl_sim* s, int64_t t, const int64_t* kbase, int k_local) {
  const DevTables& dt = s->dt;
  StepArgs sa;
  sa.nlb = s->nlb;
  sa.fblk = s->fblk;
  sa.xblocks = (s->g.nx + STEP_T - 1) / STEP_T;
  sa.sqrt_cs2 = SQRT_CS2;
  const int64_t ns2 = (int64_t)std::max(dt.nslots, 1) * 2;
  sa.fc_cur = dt.fcontrib + (t & 1) * ns2;
  sa.fc_prev = dt.fcontrib + ((t + 1) & 1) * ns2;
  sa.ctls = s->run_ctl;
  sa.kbase = kbase;
  sa.k_local = k_local;
  sa.hist = dt.ngroups > 0 ? s->run_hist : nullptr;
  const int64_t np1 = std::max(dt.np, 1);
  sa.pv_cur = dt.pvel + (t & 1) * np1;
  sa.pv_next = dt.pvel + ((t + 1) & 1) * np1;
  return sa;
}

// Joins the side stream of overlapped groups back into the main stream (before anything that
// reads or writes nodes the side stream may still be stepping).
static vl_status join_overlap(vl_sim* s) {
  if (!s->ov_open) return VL_OK;
  VL_HIP(hipEventRecord(s->ev_join, s->side));
  VL_HIP(hipStreamWaitEvent(s->stream, s->ev_join, 0));
  s->ov_open = false;
  return VL_OK;
}

// Split mode: the boundary kernel of one step on the side stream, forked from the main stream.
// The caller joins it (ev_join) before anything that reads what it wrote.
template <class S, int M>
static vl_status fork_boundary(vl_sim* s, const typename S::T* A, typename S::T* B, const StepArgs& sa) {
  VL_HIP(hipEventRecord(s->ev_fork, s->stream));
  VL_HIP(hipStreamWaitEvent(s->side, s->ev_fork, 0));
  if (sa.nlb > 0)
    hipLaunchKernelGGL((k_boundary<S, M>), dim3((unsigned)sa.nlb), dim3(STEP_T), 0, s->side, A, B, s->g, s->coll,
                       s->dt, sa);
  return VL_OK;
}

// One single step of every node a K-step pass does not cover (boundary and listed interior).
template <class S, int M>
static void launch_phase(vl_sim* s, hipStream_t st, const typename S::T* A, typename S::T* B, const StepArgs& sa) {
  const int nblk = sa.nlb + (s->nlisted + STEP_T - 1) / STEP_T;
  if (nblk > 0)
    hipLaunchKernelGGL((k_phase<S, M>), dim3((unsigned)nblk), dim3(STEP_T), 0, st, A, B, s->dlisted, s->nlisted,
                       s->g, s->coll, s->dt, sa);
}

// Launch one step.  Fused mode: one kernel.  Split mode: the boundary kernel on the side stream
// and the bulk on the main stream, forked after the previous step and joined before the next.
template <class S, int M>
static vl_status launch_step(vl_sim* s, const int64_t* kbase, int k_local, bool timed_ok) {
  using T = typename S::T;
  VL_TRY(join_overlap(s));
  const T* A = (const T*)s->buf[s->cur];
  T* B = (T*)s->buf[(s->cur + 1) % s->nbuf];
  const Geom& g = s->g;
  const StepArgs sa = step_args(s, s->t, kbase, k_local);
  const bool timed = timed_ok && s->timing && s->ev_used + 2 <= (int64_t)s->ev.size();
  if (!s->split) {
    if (timed) VL_HIP(hipEventRecord(s->ev[s->ev_used], s->stream));
    hipLaunchKernelGGL((k_step<S, M>), dim3((unsigned)(sa.nlb + sa.xblocks * g.ny)), dim3(STEP_T), 0, s->stream, A,
                       B, s->dflag, s->drows, g, s->coll, s->dt, sa);
  } else {
    if (sa.nlb > 0) {
      VL_TRY((fork_boundary<S, M>(s, A, B, sa)));
      VL_HIP(hipEventRecord(s->ev_join, s->side));
    }
    if (timed) VL_HIP(hipEventRecord(s->ev[s->ev_used], s->stream));
    const int v = s->bulk_v;
    const int xb = (g.nx + BULK_T * std::max(v, 1) - 1) / (BULK_T * std::max(v, 1));
    const dim3 grid((unsigned)(xb * g.ny));
    switch (v) {
      case 0:
        hipLaunchKernelGGL((k_bulk<S, M, 0>), grid, dim3(BULK_T), 0, s->stream, A, B, s->dflag, s->drows, g,
                           s->coll, xb);
        break;
      case 1:
        hipLaunchKernelGGL((k_bulk<S, M, 1>), grid, dim3(BULK_T), 0, s->stream, A, B, s->dflag, s->drows, g,
                           s->coll, xb);
        break;
      case 2:
        hipLaunchKernelGGL((k_bulk<S, M, 2>), grid, dim3(BULK_T), 0, s->stream, A, B, s->dflag, s->drows, g,
                           s->coll, xb);
        break;
      case 4:
        if constexpr (sizeof(T) <= 4)
          hipLaunchKernelGGL((k_bulk<S, M, 4>), grid, dim3(BULK_T), 0, s->stream, A, B, s->dflag, s->drows, g,
                           s->coll, xb);
        break;
      default:
        if constexpr (sizeof(T) == 2)
          hipLaunchKernelGGL((k_bulk<S, M, 8>), grid, dim3(BULK_T), 0, s->stream, A, B, s->dflag, s->drows, g,
                           s->coll, xb);
    }
  }
  if (timed) {
    VL_HIP(hipEventRecord(s->ev[s->ev_used + 1], s->stream));
    s->ev_used += 2;
  }
  if (s->split && sa.nlb > 0) VL_HIP(hipStreamWaitEvent(s->stream, s->ev_join, 0));
  VL_HIP(hipGetLastError());
  s->cur = (s->cur + 1) % s->nbuf;
  s->t += 1;
  return VL_OK;
}

// The pass kernel over the tasks [tasks, tasks + n) on stream `st`, of which the first nj are
// joined (whole workgroups): one launch for those, one for the rest.
template <class S, int M, int W, int V, int D, int K>
static void launch_pass_w(vl_sim* s, hipStream_t st, const int4* tasks, int n, int nj, const typename S::T* A,
                          const PassOut<typename S::T>& po) {
  if (nj > 0)
    hipLaunchKernelGGL((k_pass<S, M, W, V, D, K, true>), dim3((unsigned)(nj / W)), dim3(W * TB_LANES), 0, st, A, po,
                       tasks, nj, s->g, s->coll);
  if (n > nj)
    hipLaunchKernelGGL((k_pass<S, M, W, V, D, K, false>), dim3((unsigned)((n - nj + W - 1) / W)), dim3(W * TB_LANES),
                       0, st, A, po, tasks + nj, n - nj, s->g, s->coll);
}

template <class S, int M, int V, int D, int K>
static void launch_pass(vl_sim* s, hipStream_t st, const int4* tasks, int n, int nj, const typename S::T* A,
                        const PassOut<typename S::T>& po) {
  if (s->tb_waves == 8)
    launch_pass_w<S, M, 8, V, D, K>(s, st, tasks, n, nj, A, po);
  else
    launch_pass_w<S, M, 4, V, D, K>(s, st, tasks, n, nj, A, po);
}

template <class S, int M>
static void launch_pass_tasks(vl_sim* s, hipStream_t st, const int4* tasks, int n, int nj, const typename S::T* A,
                              const PassOut<typename S::T>& po) {
  if (n <= 0) return;
  const int K = s->tb_k;
  const bool wide = sizeof(typename S::T) < 8 && s->tb_v == 2;  // V = 2 (D = 2) or V = 1 (D = 3)
  if (K == 5) {
    if constexpr (sizeof(typename S::T) == 4)  // FP32 only (FP64 and FP16 are slower above 4 / 3)
      launch_pass<S, M, 1, 3, 5>(s, st, tasks, n, nj, A, po);
  } else if (K == 4) {
    if (wide)
      launch_pass<S, M, 2, 2, 4>(s, st, tasks, n, nj, A, po);
    else
      launch_pass<S, M, 1, 3, 4>(s, st, tasks, n, nj, A, po);
  } else if (K == 3) {
    if (wide)
      launch_pass<S, M, 2, 2, 3>(s, st, tasks, n, nj, A, po);
    else
      launch_pass<S, M, 1, 3, 3>(s, st, tasks, n, nj, A, po);
  } else {
    if (wide)
      launch_pass<S, M, 2, 2, 2>(s, st, tasks, n, nj, A, po);
    else
      launch_pass<S, M, 1, 3, 2>(s, st, tasks, n, nj, A, po);
  }
}


// K steps, t -> t+K, through the buffers b_j = buf[(cur + j) % (K + 1)] (b_j holds time t+j).
// The pass kernel takes the covered nodes from b_0 straight to b_K (and writes its ring nodes'
// values of time t+j to b_j); the single-step phases take every other fluid node through
// b_0 -> b_1 -> ... -> b_K, one launch per step.  Every node of b_K is written once.
//
// Overlapped (the default): the far tasks' pass runs on the main stream; the near tasks' pass
// and the K phases run on the side stream.
//   * near(g) waits for far(g-1) (its region holds far outputs) and, in stream order, for the
//     phases of g-1 (it holds uncovered nodes);
//   * far(g) waits for near(g-1) only: no far region holds an uncovered node, so far(g) runs
//     while the phases of g-1 finish.
// The far outputs are disjoint from the uncovered nodes and from the near outputs, and the
// phases read and write only uncovered nodes and ring nodes (near outputs), so nothing a far
// pass writes is touched by a phase still running; near(g+1) waits for far(g) before it writes
// the ring nodes that far(g) may read.  Events alternate by group parity.
template <class S, int M>
static vl_status launch_group(vl_sim* s, const int64_t* kbase, int k_local, bool timed_ok) {
  using T = typename S::T;
  const int K = s->tb_k;
  T* b[TB_MAX_K + 1];
  for (int j = 0; j <= K; ++j) b[j] = (T*)s->buf[(s->cur + j) % (K + 1)];
  PassOut<T> po{};
  for (int j = 1; j < K; ++j) po.ring[j - 1] = b[j];
  po.out = b[K];
  const bool timed = timed_ok && s->timing && s->ev_used + 2 <= (int64_t)s->ev.size();
  const int nfar = s->ntasks_far, nnear = s->ntasks - nfar;
  bool first = false;
  if (!s->ov_open) {  // fork the side stream from everything before this group
    VL_HIP(hipEventRecord(s->ev_fork, s->stream));
    VL_HIP(hipStreamWaitEvent(s->side, s->ev_fork, 0));
    s->ov_open = true;
    first = true;
  }
  const int p = s->ov_parity;
  // side stream: near pass, phase 1, (event), phases 2..K
  if (!first) VL_HIP(hipStreamWaitEvent(s->side, s->ev_far[p ^ 1], 0));
  launch_pass_tasks<S, M>(s, s->side, s->dtasks + nfar, nnear, s->tb_joined_near, b[0], po);
  if (!s->tb_overlap) launch_pass_tasks<S, M>(s, s->side, s->dtasks, nfar, s->tb_joined_far, b[0], po);
  launch_phase<S, M>(s, s->side, b[0], b[1], step_args(s, s->t, kbase, k_local));
  VL_HIP(hipEventRecord(s->ev_near[p], s->side));
  for (int j = 2; j <= K; ++j)
    launch_phase<S, M>(s, s->side, b[j - 1], b[j], step_args(s, s->t + j - 1, kbase, k_local + j - 1));
  // main stream: far pass
  if (!first) VL_HIP(hipStreamWaitEvent(s->stream, s->ev_near[p ^ 1], 0));
  if (timed) VL_HIP(hipEventRecord(s->ev[s->ev_used], s->stream));
  if (s->tb_overlap) launch_pass_tasks<S, M>(s, s->stream, s->dtasks, nfar, s->tb_joined_far, b[0], po);
  if (timed) {
    VL_HIP(hipEventRecord(s->ev[s->ev_used + 1], s->stream));
    s->ev_used += 2;
  }
  VL_HIP(hipEventRecord(s->ev_far[p], s->stream));
  s->ov_parity ^= 1;
  if (!s->tb_overlap) VL_TRY(join_overlap(s));  // not overlapped: the group ends complete
  VL_HIP(hipGetLastError());
  s->cur = (s->cur + K) % (K + 1);
  s->t += K;
  return VL_OK;
}

template <class S, int M>
static vl_status launch_steps(vl_sim* s, int n, const int64_t* kbase, int k_local, bool timed_ok) {
  return n > 1 ? launch_group<S, M>(s, kbase, k_local, timed_ok) : launch_step<S, M>(s, kbase, k_local, timed_ok);
}

template <class S>
static vl_status step_dispatch(vl_sim* s, int n, const int64_t* kbase, int k_local, bool timed_ok) {
  switch (s->coll.model) {
    case MODEL_FORCED: return launch_steps<S, MODEL_FORCED>(s, n, kbase, k_local, timed_ok);
    case MODEL_BGK: return launch_steps<S, MODEL_BGK>(s, n, kbase, k_local, timed_ok);
    default: return launch_steps<S, MODEL_TRT>(s, n, kbase, k_local, timed_ok);
  }
}

// Launches the next `n` steps (n = 1, or n = K with K-step passes) as steps k_local.. of a window.
static vl_status some_steps(vl_sim* s, int n, const int64_t* kbase, int k_local, bool timed_ok) {
  switch (s->precision) {
    case VL_F32: return step_dispatch<StoreF32>(s, n, kbase, k_local, timed_ok);
    case VL_F16: return step_dispatch<StoreF16>(s, n, kbase, k_local, timed_ok);
    default: return step_dispatch<StoreF64>(s, n, kbase, k_local, timed_ok);
  }
}

// Steps advanced per launch group when that many remain.
static int group_steps(const vl_sim* s) { return s->tb_k; }

// Outward velocities of the outlet-controller nodes from the current A (start of a run).
static vl_status prime_outlet(vl_sim* s) {
  DevTables& t = s->dt;
  double* pv = t.pvel + (s->t & 1) * (int64_t)std::max(t.np, 1);
  const unsigned grid = (unsigned)((t.np + 255) / 256);
  switch (s->precision) {
    case VL_F32:
      hipLaunchKernelGGL(k_outlet_prime<StoreF32>, dim3(grid), dim3(256), 0, s->stream,
                         (const float*)s->buf[s->cur], s->g, s->coll, t, pv);
      break;
    case VL_F16:
      hipLaunchKernelGGL(k_outlet_prime<StoreF16>, dim3(grid), dim3(256), 0, s->stream,
                         (const _Float16*)s->buf[s->cur], s->g, s->coll, t, pv);
      break;
    default:
      hipLaunchKernelGGL(k_outlet_prime<StoreF64>, dim3(grid), dim3(256), 0, s->stream,
                         (const double*)s->buf[s->cur], s->g, s->coll, t, pv);
  }
  VL_HIP(hipGetLastError());
  return VL_OK;
}

// Group forces of the step that just ran (time s->t - 1) into `out`.
static vl_status finish_forces(vl_sim* s, double* out) {
  DevTables& t = s->dt;
  if (t.ngroups == 0) return VL_OK;
  const int64_t ns2 = (int64_t)std::max(t.nslots, 1) * 2;
  const double* fc = t.fcontrib + ((s->t - 1) & 1) * ns2;
  hipLaunchKernelGGL(k_force, dim3(1), dim3(STEP_T), 0, s->stream, fc, t.gfirst, t.ngroups, out);
  VL_HIP(hipGetLastError());
  return VL_OK;
}

template <class V>
static vl_status upload(vl_sim* s, V*& dst, const std::vector<V>& src) {
  if (src.empty()) return VL_OK;
  const size_t bytes = src.size() * sizeof(V);
  VL_HIP(hipMalloc((void**)&dst, bytes));
  s->device_bytes += (int64_t)bytes;
  VL_HIP(hipMemcpy(dst, src.data(), bytes, hipMemcpyHostToDevice));
  return VL_OK;
}


static vl_status ensure_hist(vl_sim* s, int64_t steps) {
  if (steps <= s->hist_cap) return VL_OK;
  if (s->dhist) VL_HIP(hipFree(s->dhist));
  s->dhist = nullptr;
  const int64_t ng = std::max<int64_t>(s->dt.ngroups, 1);
  VL_HIP(hipMalloc((void**)&s->dhist, (size_t)(steps * ng * 2) * sizeof(double)));
  s->hist_cap = steps;
  return VL_OK;
}

// Equilibrium populations of (rho, ux, uy) stored in the sim's precision at node (x, y).
static void host_store_eq(vl_sim* s, std::vector<char>& host, int64_t x, int64_t y, double rho, double ux,
                          double uy) {
  double feq[Q];
  Exact::equilibrium(rho, ux, uy, feq);
  const int64_t off = y * s->g.pitch + x;
  for (int i = 0; i < Q; ++i) {
    const int64_t e = i * s->g.plane + off;
    if (s->precision == VL_F64)
      ((double*)host.data())[e] = feq[i];
    else if (s->precision == VL_F32)
      ((float*)host.data())[e] = (float)(feq[i] - wgt(i));
    else
      ((_Float16*)host.data())[e] = (_Float16)((float)(feq[i] - wgt(i)) * F16_SCALE);
  }
}

static double host_load(const vl_sim* s, const std::vector<char>& host, int i, int64_t off) {
  const int64_t e = i * s->g.plane + off;
  if (s->precision == VL_F64) return ((const double*)host.data())[e];
  if (s->precision == VL_F32) return wgt(i) + (double)((const float*)host.data())[e];
  return wgt(i) + (double)((float)((const _Float16*)host.data())[e] * F16_INV);
}

static vl_status read_raw(vl_sim* s, std::vector<char>& host) {
  host.resize(s->pop_bytes);
  VL_HIP(hipStreamSynchronize(s->stream));
  VL_HIP(hipMemcpy(host.data(), s->buf[s->cur], s->pop_bytes, hipMemcpyDeviceToHost));
  return VL_OK;
}

static vl_status write_raw(vl_sim* s, const std::vector<char>& host) {
  VL_HIP(hipStreamSynchronize(s->stream));
  for (int b = 0; b < s->nbuf; ++b) VL_HIP(hipMemcpy(s->buf[b], host.data(), s->pop_bytes, hipMemcpyHostToDevice));
  s->cur = 0;
  return VL_OK;
}

static void free_sim(vl_sim* s) {
  if (!s) return;
  if (s->stream) (void)hipStreamSynchronize(s->stream);
  auto f = [](void* p) {
    if (p) (void)hipFree(p);
  };
  for (void* b : s->buf) f(b);
  f(s->dtasks);
  f(s->dlisted);
  f(s->dflag);
  f(s->drows);
  DevTables& t = s->dt;
  for (void* p : {(void*)t.bx, (void*)t.by, (void*)t.bfirst, (void*)t.dir, (void*)t.kind, (void*)t.vgroup,
                  (void*)t.lflags, (void*)t.lnx, (void*)t.lny, (void*)t.fslot, (void*)t.q, (void*)t.uwx,
                  (void*)t.uwy, (void*)t.cu6, (void*)t.rho, (void*)t.px, (void*)t.py, (void*)t.pnx,
                  (void*)t.pny, (void*)t.gfirst, (void*)t.fcontrib, (void*)t.drho, (void*)t.pvel})
    f(p);
  f(s->dhist);
  f(s->d_ctl);
  f(s->d_kbase);
  if (s->h_ctl) (void)hipHostFree(s->h_ctl);
  for (auto& kv : s->graphs)
    if (kv.second.first) (void)hipGraphExecDestroy(kv.second.first);
  for (hipEvent_t e : s->ev) (void)hipEventDestroy(e);
  for (hipEvent_t e : s->ev_chunk)
    if (e) (void)hipEventDestroy(e);
  if (s->side) (void)hipStreamSynchronize(s->side);
  if (s->ev_fork) (void)hipEventDestroy(s->ev_fork);
  if (s->ev_join) (void)hipEventDestroy(s->ev_join);
  for (hipEvent_t e : {s->ev_far[0], s->ev_far[1], s->ev_near[0], s->ev_near[1]})
    if (e) (void)hipEventDestroy(e);
  if (s->side) (void)hipStreamDestroy(s->side);
  if (s->stream) (void)hipStreamDestroy(s->stream);
  delete s;
}

static int64_t round_up(int64_t v, int64_t m) { return (v + m - 1) / m * m; }

// Task list of K-step passes, from the node flags (0 interior, 1 boundary, 2 solid):
//  * strips of up to tb_out(K, V) output columns (a multiple of V) start at x = x0,
//    x0 + tb_out(K, V), ... (V columns per lane);
//  * with `join` = W > 1, the columns from x0 first take as many joined strips as fit, each
//    swept by a workgroup of W waves together (64 V W columns, of which tb_halo(K, V) on the
//    left and K on the right are margin); then single-wave strips follow on the grid, the
//    first cut short at the joined strips' end;
//  * a row is usable for a strip when every node of the strip's columns plus K - 1 on each side
//    is interior and its pulls stay inside the domain without wrapping (rows 1 .. ny-2,
//    columns 1 .. nx-2);
//  * a run of usable rows [ya, yb] gives output rows ya+K-1 .. yb-K+1 (the margin of the levels
//    below K), cut into tasks of at most `rows` rows of near-equal size;
//  * the rows of a joined strip that its own runs leave out (where a body or wall crosses its
//    wide footprint) go to single-wave strips over its columns (the grid's strips, cut at the
//    joined strip's ends), with their own runs; so a joined plan covers every node the
//    single-wave plan covers.
// A joined task is W wave tasks in a row (bits 12 and 13 set, each wave's output columns and
// its offset in its loaded columns in bits 16..23).  Single-wave tasks are ordered by (first
// row, first column) so a workgroup's waves sweep adjacent strips; joined groups likewise among
// themselves.  Every interior node no task covers goes to the single-step list; a covered node
// is a ring node when one of its eight neighbours is a fluid node that no task covers, and a
// task marks each of its four edges that holds a ring node.  Order: far joined groups, far
// single-wave tasks, near joined groups, near single-wave tasks (far and near are launched as
// separate ranges, so each range starts with its joined groups, whole workgroups from its first
// task).  The caller sets the lockstep bits of the single-wave tasks, which depend on how it
// groups tasks into workgroups.
vl_status plan_passes(const std::vector<uint8_t>& fl, int nx, int ny, int64_t pitch, int K, int V, int rows,
                      int x0_grid, int join, PassPlan& plan) {
  if (nx >= 65536 || ny >= 65536 || pitch * ny >= ((int64_t)1 << 31) || rows < 1 || K < 2 || K > TB_MAX_K ||
      (V != 1 && V != 2) || join < 0 || join > 16)
    return set_error(VL_EINVAL, "K-step passes need nx, ny < 65536, a plane below 2^31 nodes and rows >= 1");
  std::vector<int4> tasks, groups;  // single-wave tasks; joined groups (join tasks each)
  std::vector<uint8_t> covered((size_t)nx * ny, 0);
  std::vector<uint8_t> usable((size_t)ny), joined_row((size_t)ny);
  const int halo = tb_halo(K, V), out = tb_out(K, V);
  const int gout = join > 1 ? (TB_LANES * V * join - halo - K) / V * V : 0;
  // Columns interior in at least half the rows: the strips' footprints stay within
  // [xa, xb], so that a column of boundary nodes at an inlet or outlet does not disqualify a
  // whole strip.  The strips are the aligned grid from tb_x0, a narrower first strip ending at
  // tb_x0, and a last strip clipped at xb.
  int xa = 1, xb = nx - 2;
  {
    std::vector<int> interior((size_t)nx, 0);
    for (int y = 0; y < ny; ++y)
      for (int x = 0; x < nx; ++x) interior[(size_t)x] += fl[(size_t)y * nx + x] == 0;
    while (xa < nx - 1 && 2 * interior[(size_t)xa] < ny) ++xa;
    while (xb > 0 && 2 * interior[(size_t)xb] < ny) --xb;
    xa = std::max(xa, 1);
    xb = std::min(xb, nx - 2);
  }
  // the next line of the single-wave grid (x0_grid + a multiple of tb_out) after column x
  auto grid_end = [&](int x) { return x0_grid + ((x - x0_grid) / out + 1) * out; };
  std::vector<std::tuple<int, int, bool>> strips;  // (first output column, output columns, joined)
  {
    int x_first = std::max(halo, xa + K - 1);
    x_first += (V - (x_first - halo) % V) % V;  // the strip's first column is a multiple of V
    if (x_first < x0_grid) {
      const int w = std::min({out, x0_grid - x_first, xb - (K - 1) - x_first + 1}) / V * V;
      if (w >= 1) strips.emplace_back(x_first, w, false);
    }
    int x0 = std::max(x0_grid, x_first);
    if (gout > 0)
      for (; x0 + gout <= std::min(nx - K, xb - (K - 1) + 1); x0 += gout) strips.emplace_back(x0, gout, true);
    // single-wave strips on the grid from x0_grid (the first one ends at the next grid line)
    while (x0 < nx) {
      const int w = std::min({grid_end(x0) - x0, nx - K - x0, xb - (K - 1) - x0 + 1}) / V * V;
      if (w < 1) break;
      strips.emplace_back(x0, w, false);
      x0 += w;
    }
  }
  // The output row segments of a strip: usable runs less their margins, less the rows in
  // `skip` (if any), cut into pieces of at most `rows`.
  auto segments = [&](int x0, int w, const std::vector<uint8_t>* skip, auto&& emit) {
    for (int y = 0; y < ny; ++y) {
      bool ok = y >= 1 && y <= ny - 2;
      for (int x = x0 - (K - 1); ok && x <= x0 + w + K - 2; ++x) ok = fl[(size_t)y * nx + x] == 0;
      usable[(size_t)y] = ok;
    }
    for (int y = 0; y < ny;) {
      if (!usable[(size_t)y]) {
        ++y;
        continue;
      }
      int yb = y;
      while (yb + 1 < ny && usable[(size_t)yb + 1]) ++yb;
      for (int a = y + K - 1, end = yb - K + 2; a < end;) {  // output rows [y + K - 1, yb - K + 1]
        if (skip && (*skip)[(size_t)a]) {
          ++a;
          continue;
        }
        int b = a;
        while (b < end && !(skip && (*skip)[(size_t)b])) ++b;
        const int len = b - a, nseg = (len + rows - 1) / rows;
        for (int k = 0; k < nseg; ++k)
          emit(a + (int)((int64_t)len * k / nseg), a + (int)((int64_t)len * (k + 1) / nseg));
        a = b;
      }
      y = yb + 1;
    }
  };
  auto add_task = [&](int x0, int ya, int yb, int w) {
    tasks.push_back(make_int4(x0, ya, yb - ya, w));
    for (int yy = ya; yy < yb; ++yy) std::memset(&covered[(size_t)yy * nx + x0], 1, (size_t)w);
  };
  for (const auto& [x0, w, joined] : strips) {
    if (!joined) {
      segments(x0, w, nullptr, [&](int ya, int yb) { add_task(x0, ya, yb, w); });
      continue;
    }
    std::fill(joined_row.begin(), joined_row.end(), 0);
    const int xs = x0 - halo;  // the group's first loaded column
    segments(x0, w, nullptr, [&](int ya, int yb) {
      for (int j = 0; j < join; ++j) {
        const int lx = xs + TB_LANES * V * j;
        const int ox0 = std::max(x0, lx), ox1 = std::min(x0 + w, lx + TB_LANES * V);
        groups.push_back(make_int4(ox0, ya, yb - ya, (ox1 - ox0) | 3 << 12 | (ox0 - lx) << 16));
        for (int yy = ya; yy < yb; ++yy) std::memset(&covered[(size_t)yy * nx + ox0], 1, (size_t)(ox1 - ox0));
      }
      std::fill(joined_row.begin() + ya, joined_row.begin() + yb, 1);
    });
    // the rows it leaves, in single-wave strips on the grid cut at its ends (so they cover at
    // least what the grid's strips would)
    for (int nx0 = x0; nx0 < x0 + w;) {
      const int nw = std::min(grid_end(nx0), x0 + w) - nx0;
      segments(nx0, nw, &joined_row, [&](int ya, int yb) { add_task(nx0, ya, yb, nw); });
      nx0 += nw;
    }
  }
  std::sort(tasks.begin(), tasks.end(),
            [](const int4& a, const int4& b) { return a.y != b.y ? a.y < b.y : a.x < b.x; });
  const int ngroups = join > 1 ? (int)(groups.size() / (size_t)join) : 0;
  {
    std::vector<int> order((size_t)ngroups);
    for (int i = 0; i < ngroups; ++i) order[(size_t)i] = i;
    std::sort(order.begin(), order.end(), [&](int a, int b) {
      const int4 &ta = groups[(size_t)a * join], &tb = groups[(size_t)b * join];
      return ta.y != tb.y ? ta.y < tb.y : ta.x < tb.x;
    });
    std::vector<int4> sorted;
    sorted.reserve(groups.size());
    for (int i : order)
      for (int j = 0; j < join; ++j) sorted.push_back(groups[(size_t)i * join + j]);
    groups.swap(sorted);
  }
  std::vector<uint32_t>& listed = plan.listed;
  listed.clear();
  std::vector<uint8_t> ring((size_t)nx * ny, 0);
  int64_t ncov = 0;
  for (int y = 0; y < ny; ++y)
    for (int x = 0; x < nx; ++x) {
      const size_t k = (size_t)y * nx + x;
      if (fl[k] != 0) continue;
      if (!covered[k]) {
        listed.push_back((uint32_t)y << 16 | (uint32_t)x);
        continue;
      }
      ++ncov;
      bool r = false;  // covered nodes lie at least one node inside the domain edges
      for (int dy = -1; dy <= 1 && !r; ++dy)
        for (int dx = -1; dx <= 1 && !r; ++dx) {
          const size_t kn = (size_t)(y + dy) * nx + (x + dx);
          r = fl[kn] != 2 && !covered[kn];
        }
      ring[k] = r;
    }
  for (std::vector<int4>* list : {&tasks, &groups})
    for (int4& tk : *list) {
      const int xa = tk.x, xb = tk.x + (tk.w & 0xff) - 1, ya = tk.y, yb = tk.y + tk.z - 1;
      int edges = 0;
      for (int yy = ya; yy <= yb; ++yy)
        for (int xx = xa; xx <= xb; ++xx) {
          if (!ring[(size_t)yy * nx + xx]) continue;
          if (yy != ya && yy != yb && xx != xa && xx != xb)
            return set_error(VL_EINVAL, "internal: a ring node inside a pass task");
          edges |= (yy == ya ? 1 : 0) | (yy == yb ? 2 : 0) | (xx == xa ? 4 : 0) | (xx == xb ? 8 : 0);
        }
      tk.w |= edges << 8;
    }
  // Far and near tasks.  A task is near when the region its outputs depend on (K nodes around
  // them) holds a fluid node that no task covers: its pass must wait for the single steps of
  // those nodes.  Far tasks depend on covered nodes only, so the next group's far pass can run
  // while the single steps of this group's uncovered nodes go on.  Far tasks come first in the
  // list; ring nodes are all in near tasks.
  std::vector<int64_t> unc((size_t)(nx + 1) * (ny + 1), 0);  // 2-D prefix count of uncovered fluid
  for (int y = 0; y < ny; ++y)
    for (int x = 0; x < nx; ++x) {
      const size_t k = (size_t)y * nx + x;
      unc[(size_t)(y + 1) * (nx + 1) + x + 1] = unc[(size_t)y * (nx + 1) + x + 1] +
                                                unc[(size_t)(y + 1) * (nx + 1) + x] -
                                                unc[(size_t)y * (nx + 1) + x] + (fl[k] != 2 && !covered[k]);
    }
  auto uncovered_in = [&](int xa, int ya, int xb, int yb) {  // inclusive, clipped to the domain
    xa = std::max(xa, 0), ya = std::max(ya, 0), xb = std::min(xb, nx - 1), yb = std::min(yb, ny - 1);
    return unc[(size_t)(yb + 1) * (nx + 1) + xb + 1] - unc[(size_t)ya * (nx + 1) + xb + 1] -
           unc[(size_t)(yb + 1) * (nx + 1) + xa] + unc[(size_t)ya * (nx + 1) + xa];
  };
  auto is_far = [&](const int4& tk) {
    const int w = tk.w & 0xff;
    return uncovered_in(tk.x - K, tk.y - K, tk.x + w - 1 + K, tk.y + tk.z - 1 + K) == 0;
  };
  std::stable_partition(tasks.begin(), tasks.end(), is_far);
  // a joined group is far when all its waves are
  std::vector<int4> gfar, gnear;
  for (int g = 0; g < ngroups; ++g) {
    bool far = true;
    for (int j = 0; j < join; ++j) far = far && is_far(groups[(size_t)g * join + j]);
    std::vector<int4>& dst = far ? gfar : gnear;
    dst.insert(dst.end(), groups.begin() + (int64_t)g * join, groups.begin() + (int64_t)(g + 1) * join);
  }
  int nfar1 = 0;  // far single-wave tasks
  while (nfar1 < (int)tasks.size() && is_far(tasks[(size_t)nfar1])) ++nfar1;
  std::vector<int4>& all = plan.tasks;
  all.clear();
  all.insert(all.end(), gfar.begin(), gfar.end());
  all.insert(all.end(), tasks.begin(), tasks.begin() + nfar1);
  all.insert(all.end(), gnear.begin(), gnear.end());
  all.insert(all.end(), tasks.begin() + nfar1, tasks.end());
  const int nfar = (int)gfar.size() + nfar1;
  int64_t far_nodes = 0;
  for (int i = 0; i < nfar; ++i) {
    const int4& tk = all[(size_t)i];
    if ((tk.w >> 8) & 15) return set_error(VL_EINVAL, "internal: a ring node in a far task");
    far_nodes += (int64_t)(tk.w & 0xff) * tk.z;
  }
  plan.nfar = nfar;
  plan.njoined_far = (int)gfar.size();
  plan.njoined_near = (int)gnear.size();
  plan.far_nodes = far_nodes;
  plan.covered = ncov;
  return VL_OK;
}

static bool lockstep_group_ok(const std::vector<int4>& tasks, const int32_t* owner, int64_t g0, int W) {
  for (int64_t i = g0 + 1; i < g0 + W; ++i)
    if (tasks[(size_t)i].y != tasks[(size_t)g0].y || tasks[(size_t)i].z != tasks[(size_t)g0].z ||
        (owner && owner[i] != owner[g0]))
      return false;
  return true;
}

static bool is_joined(const int4& t) { return (t.w >> 13) & 1; }

void mark_lockstep(std::vector<int4>& tasks, const int32_t* owner, const TaskRanges& ranges, int W) {
  for (int4& t : tasks)
    if (!is_joined(t)) t.w &= ~(1 << 12);
  for (const auto& [begin, end] : ranges)
    for (int64_t g0 = begin; g0 + W <= end; g0 += W) {
      bool any_joined = false;
      for (int64_t i = g0; i < g0 + W; ++i) any_joined = any_joined || is_joined(tasks[(size_t)i]);
      if (!any_joined && lockstep_group_ok(tasks, owner, g0, W))
        for (int64_t i = g0; i < g0 + W; ++i) tasks[(size_t)i].w |= 1 << 12;
    }
}

vl_status check_lockstep(const std::vector<int4>& tasks, const int32_t* owner, const TaskRanges& ranges, int W,
                         int wave_cols) {
  if (W < 1) return set_error(VL_EINVAL, "internal: lockstep workgroups of no waves");
  std::vector<uint8_t> launched(tasks.size(), 0);
  for (const auto& [begin, end] : ranges) {
    if (begin < 0 || end > (int64_t)tasks.size() || begin > end)
      return set_error(VL_EINVAL, "internal: a pass launch range outside the task list");
    for (int64_t g0 = begin; g0 < end; g0 += W) {
      const int64_t g1 = std::min<int64_t>(g0 + W, end);
      int marked = 0, joined = 0;
      for (int64_t i = g0; i < g1; ++i) {
        marked += (tasks[(size_t)i].w >> 12) & 1;
        joined += is_joined(tasks[(size_t)i]);
        launched[(size_t)i] = 1;
      }
      if (marked == 0 && joined == 0) continue;
      if (marked != W || g1 - g0 != W || !lockstep_group_ok(tasks, owner, g0, W))
        return set_error(VL_EINVAL, "internal: a lockstep workgroup whose waves would meet different barriers");
      if (joined == 0) continue;
      if (joined != W || wave_cols == 0)
        return set_error(VL_EINVAL, "internal: joined pass tasks that do not fill a workgroup");
      // wave i loads the wave_cols columns after wave i - 1's: first loaded column = first
      // output column less the offset in bits 16..23 (wave_cols < 0: 64 or 128, the same for
      // every wave of the group)
      auto loaded = [](const int4& t) { return t.x - ((t.w >> 16) & 0xff); };
      const int step = wave_cols > 0 ? wave_cols : loaded(tasks[(size_t)g0 + 1]) - loaded(tasks[(size_t)g0]);
      if (step != TB_LANES && step != 2 * TB_LANES)
        return set_error(VL_EINVAL, "internal: joined pass tasks whose waves are not adjacent strips");
      for (int64_t i = g0 + 1; i < g1; ++i) {
        const int4 &a = tasks[(size_t)i - 1], &b = tasks[(size_t)i];
        if (loaded(b) - loaded(a) != step || b.x != a.x + (a.w & 0xff))
          return set_error(VL_EINVAL, "internal: joined pass tasks whose waves are not adjacent strips");
      }
    }
  }
  for (size_t i = 0; i < tasks.size(); ++i)
    if (((tasks[i].w >> 12) & 1) && !launched[i])
      return set_error(VL_EINVAL, "internal: a lockstep task outside every launch range");
  return VL_OK;
}

// The launch ranges of a domain's task list: far tasks, then near tasks (two launches, on two
// streams when the far pass overlaps the single steps, else one after the other).
static TaskRanges pass_ranges(int64_t ntasks, int64_t nfar) { return {{0, nfar}, {nfar, ntasks}}; }

static vl_status build_passes(vl_sim* s, int rows) {
  PassPlan plan;
  VL_TRY(plan_passes(s->tab.flag, s->g.nx, s->g.ny, s->g.pitch, s->tb_k, s->tb_v, rows, s->tb_x0,
                     s->tb_join ? s->tb_waves : 0, plan));
  std::vector<int4>& tasks = plan.tasks;
  const int nfar = plan.nfar;
  s->tb_far_nodes = plan.far_nodes;
  s->tb_joined_far = plan.njoined_far;
  s->tb_joined_near = plan.njoined_near;
  // Lockstep bit, over the ranges that launch_group launches (far and near tasks separately).
  const TaskRanges ranges = pass_ranges((int64_t)tasks.size(), nfar);
  if (s->tb_sync) mark_lockstep(tasks, nullptr, ranges, s->tb_waves);
  VL_TRY(check_lockstep(tasks, nullptr, ranges, s->tb_waves, TB_LANES * s->tb_v));
  s->tb_rows = rows;
  s->ntasks = (int)tasks.size();
  s->ntasks_far = nfar;
  s->nlisted = (int)plan.listed.size();
  s->tb_nodes = plan.covered;
  VL_TRY(upload(s, s->dtasks, tasks));
  VL_TRY(upload(s, s->dlisted, plan.listed));
  return VL_OK;
}

// Scratch area of vl_sim_run_noalloc: the per-step control table and the force-history staging
// of one call, in place of the buffers a handle allocates for vl_sim_run.  One copy per GPU,
// shared by every handle on it and held until the process ends; a call leaves its contents
// behind, and calls that overlap on one GPU race.  Host code never passes `scratch` itself to a
// kernel or a copy (on the host the name is only a stand-in for the symbol): it uses the device
// address that noalloc_scratch_setup takes with hipGetSymbolAddress, once per device.
constexpr int64_t NOALLOC_MAX_STEPS = 4096;
constexpr int64_t NOALLOC_MAX_GROUPS = 4;
constexpr size_t NOALLOC_CTL_BYTES = (NOALLOC_MAX_STEPS * sizeof(StepCtl) + 255) / 256 * 256;
constexpr size_t NOALLOC_HIST_BYTES = NOALLOC_MAX_STEPS * NOALLOC_MAX_GROUPS * 2 * sizeof(double);
static __device__ __align__(256) char scratch[NOALLOC_CTL_BYTES + NOALLOC_HIST_BYTES];
static_assert(NOALLOC_MAX_STEPS <= RUN_CHUNK, "the pinned control staging holds RUN_CHUNK steps");

static std::mutex g_scratch_mu;
static std::vector<char*> g_scratch_ptr;  // device address of `scratch`, per device

// Takes the current device's address of the scratch area (set-up of a simulation on it).
static vl_status noalloc_scratch_setup(int device) {
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if ((int)g_scratch_ptr.size() <= device) g_scratch_ptr.resize((size_t)device + 1, nullptr);
  if (g_scratch_ptr[(size_t)device]) return VL_OK;
  void* p = nullptr;
  VL_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
  if (!p) return set_error(VL_EDEVICE, "hipGetSymbolAddress returned no address for the scratch area");
  g_scratch_ptr[(size_t)device] = (char*)p;
  return VL_OK;
}

static vl_status noalloc_scratch(int device, char** out) {
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (device < 0 || (size_t)device >= g_scratch_ptr.size() || !g_scratch_ptr[(size_t)device])
    return set_error(VL_EDEVICE, "internal: the scratch area of this device was not set up");
  *out = g_scratch_ptr[(size_t)device];
  return VL_OK;
}

static vl_status create(vl::Domain& d, int precision, vl_sim** out) {
  if (precision != VL_F64 && precision != VL_F32 && precision != VL_F16)
    return set_error(VL_EINVAL, "precision must be VL_F64, VL_F32 or VL_F16");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return set_error(VL_EDEVICE, "no HIP device");
  auto* s = new vl_sim;
  std::unique_ptr<vl_sim, void (*)(vl_sim*)> guard(s, free_sim);
  VL_TRY(build_tables(d, s->tab));
  s->precision = precision;
  s->collision = d.collision;
  s->accel[0] = d.accel[0];
  s->accel[1] = d.accel[1];
  double op, om;
  omegas(d.collision, d.magic, d.tau, op, om);
  const bool forced = d.accel[0] != 0.0 || d.accel[1] != 0.0;  // src/sim.rs:395
  s->coll = Coll{op, om, d.accel[0], d.accel[1], (float)op, (float)om, (float)d.accel[0], (float)d.accel[1],
                 forced ? MODEL_FORCED : (d.collision == VL_BGK ? MODEL_BGK : MODEL_TRT)};
  s->elem = precision == VL_F64 ? 8 : (precision == VL_F32 ? 4 : 2);
  s->g.nx = (int32_t)d.nx;
  s->g.ny = (int32_t)d.ny;
  s->g.pitch = (int32_t)round_up(d.nx, 64);
  s->g.periodic_x = d.periodic_x;
  s->g.periodic_y = d.periodic_y;
  // A gap of 5 x 512 B between planes so the nine streams do not start on the same HBM channel.
  s->g.plane = round_up((int64_t)s->g.pitch * d.ny, 256) + 2560 / (int64_t)s->elem;
  s->pop_bytes = (size_t)(Q * s->g.plane) * s->elem;
  VL_HIP(hipGetDevice(&s->device));
  VL_TRY(noalloc_scratch_setup(s->device));
  VL_HIP(hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking));
  VL_HIP(hipMalloc(&s->buf[0], s->pop_bytes));
  VL_HIP(hipMalloc(&s->buf[1], s->pop_bytes));
  s->device_bytes += 2 * (int64_t)s->pop_bytes;
  // flags in pitch layout (padding columns are "not fluid")
  std::vector<uint8_t> pflag((size_t)s->g.pitch * d.ny, 2);
  for (int64_t y = 0; y < d.ny; ++y)
    std::memcpy(&pflag[y * s->g.pitch], &s->tab.flag[y * d.nx], (size_t)d.nx);
  VL_TRY(upload(s, s->dflag, pflag));
  // interior span of every row (see row_span)
  {
    std::vector<int2> rows((size_t)d.ny);
    for (int64_t y = 0; y < d.ny; ++y) {
      const uint8_t* fr = &s->tab.flag[y * d.nx];
      int lo = 0, hi = -1;
      for (int64_t x = 0; x < d.nx; ++x)
        if (fr[x] == 0) {
          if (hi < lo) lo = (int)x;
          hi = (int)x;
        }
      bool mixed = false;
      for (int x = lo; x <= hi && !mixed; ++x) mixed = fr[x] != 0;
      rows[(size_t)y] = make_int2(mixed ? -lo - 1 : lo, hi);
    }
    VL_TRY(upload(s, s->drows, rows));
  }
  // Outlet-controller nodes first: only the leading boundary workgroups need the offset.
  Tables tb = s->tab;
  s->dt.npb = outlet_nodes_first(tb);
  DevTables& t = s->dt;
  t.nb = (int32_t)tb.bnode_x.size();
  t.nl = (int32_t)tb.l_dir.size();
  t.np = (int32_t)tb.p_x.size();
  t.nslots = (int32_t)tb.f_link.size();
  t.ngroups = tb.force_groups;
  VL_TRY(upload(s, t.bx, tb.bnode_x));
  VL_TRY(upload(s, t.by, tb.bnode_y));
  VL_TRY(upload(s, t.bfirst, tb.bnode_first));
  VL_TRY(upload(s, t.dir, tb.l_dir));
  VL_TRY(upload(s, t.kind, tb.l_kind));
  VL_TRY(upload(s, t.vgroup, tb.l_vgroup));
  VL_TRY(upload(s, t.lflags, tb.l_flags));
  VL_TRY(upload(s, t.lnx, tb.l_nx));
  VL_TRY(upload(s, t.lny, tb.l_ny));
  VL_TRY(upload(s, t.fslot, tb.l_fslot));
  VL_TRY(upload(s, t.q, tb.l_q));
  VL_TRY(upload(s, t.uwx, tb.l_uwx));
  VL_TRY(upload(s, t.uwy, tb.l_uwy));
  VL_TRY(upload(s, t.cu6, tb.l_cu6));
  VL_TRY(upload(s, t.rho, tb.l_rho));
  VL_TRY(upload(s, t.px, tb.p_x));
  VL_TRY(upload(s, t.py, tb.p_y));
  VL_TRY(upload(s, t.pnx, tb.p_nx));
  VL_TRY(upload(s, t.pny, tb.p_ny));
  VL_TRY(upload(s, t.gfirst, tb.f_group_first));
  {
    const size_t fcb = (size_t)std::max(t.nslots, 1) * 4 * sizeof(double);  // two step buffers
    VL_HIP(hipMalloc((void**)&t.fcontrib, fcb));
    VL_HIP(hipMemset(t.fcontrib, 0, fcb));
  }
  s->nlb = (t.nb + STEP_T - 1) / STEP_T;
  s->fblk = t.ngroups > 0 ? s->nlb++ : -1;
  // Split the step into two concurrent kernels once the bulk is large enough for the
  // per-step fork/join to be negligible (VL_STEP_MODE=fused|split overrides).
  s->split = (int64_t)d.nx * d.ny >= ((int64_t)1 << 21);
  if (const char* e = std::getenv("VL_STEP_MODE")) s->split = std::string(e) == "split";
  // nodes per thread of the split-mode bulk (VL_BULK_V = 1, 2, 4 or 8 overrides; 4 not for
  // FP64, 8 only for FP16; 0 = one node per thread with LDS-staged x neighbours).  Measured on MI355X at 8192 x 4096: FP32 and FP64 are fastest with
  // one node per thread (4- and 8-byte loads; wider vectors cost occupancy), FP16 with two
  // (4-byte loads instead of 2-byte ones).
  s->bulk_v = precision == VL_F16 ? 2 : 1;
  if (const char* e = std::getenv("VL_BULK_V")) {
    const int v = std::atoi(e);
    if (v == 0 || v == 1 || v == 2 || (v == 4 && precision != VL_F64) || (v == 8 && precision == VL_F16))
      s->bulk_v = v;
  }
  // K-step passes on split-mode domains (VL_PASS_STEPS=1..4, 5 for FP32, overrides; above 1 it
  // also forces split mode).
  // Steps per pass and rows per task, measured on MI355X (8192 x 4096, FP16 16384 x 8192):
  // FP32 5 steps, 20 rows (16-40 rows: 1-6 % slower); FP64 4 steps, 24 rows (16: 0.5 % slower;
  // 5 steps would need more registers than its occupancy allows); FP16 3 steps, 40 rows (24:
  // 3 % slower), 2 columns per lane (its conversions make 4 steps slower).
  int k_steps = s->split ? (precision == VL_F16 ? 3 : (precision == VL_F32 ? 5 : 4)) : 1;
  if (const char* e = std::getenv("VL_PASS_STEPS")) {
    k_steps = std::min(std::max(std::atoi(e), 1), precision == VL_F32 ? TB_MAX_K : 4);
    s->split = s->split || k_steps > 1;
  }
  if (k_steps > 1) {
    s->tb_k = k_steps;
    // Layout (measured as above): 8 waves per workgroup sweeping in step (FP64: 4), strips on a
    // grid from x = 32.
    // VL_TB_ROWS, VL_TB_WAVES (4 | 8), VL_TB_SYNC (0 | 1), VL_TB_X0 (>= the halo) and VL_TB_V
    // (columns per lane, 1 | 2; FP32 / FP16 only) override.
    int rows = precision == VL_F16 ? 40 : (precision == VL_F32 ? 20 : 24);
    if (const char* e = std::getenv("VL_TB_ROWS")) rows = std::atoi(e);
    // FP64 needs about 170 VGPRs: 4-wave workgroups fit 3 to a CU, 8-wave ones only 1
    s->tb_waves = precision == VL_F64 ? 4 : 8;
    if (const char* e = std::getenv("VL_TB_WAVES")) s->tb_waves = std::atoi(e) == 4 ? 4 : 8;
    if (const char* e = std::getenv("VL_TB_SYNC")) s->tb_sync = e[0] != '0';
    if (const char* e = std::getenv("VL_TB_OVERLAP")) s->tb_overlap = e[0] != '0';
    // Joined strips (VL_TB_JOIN=1): a workgroup's waves sweep one strip and hand each other
    // their edge values through LDS, one barrier a level, instead of each recomputing a halo.
    // Bit-identical, but slower on MI355X at every layout measured (8192 x 4096 FP32 far pass
    // 835 against 743 us, FP64 1459 against 1298, FP16 16384 x 8192 2044 against 1760): the
    // barriers stall more than the halo costs.  Off by default.
    if (const char* e = std::getenv("VL_TB_JOIN")) s->tb_join = e[0] != '0';
    s->tb_v = precision == VL_F16 ? 2 : 1;
    if (const char* e = std::getenv("VL_TB_V"))
      if (precision != VL_F64) s->tb_v = std::atoi(e) == 2 ? 2 : 1;
    if (s->tb_k == 5) s->tb_v = 1;  // the 5-step pass is built with one column per lane only
    const int halo = tb_halo(s->tb_k, s->tb_v);
    s->tb_x0 = std::max(s->tb_x0, halo);
    if (const char* e = std::getenv("VL_TB_X0")) s->tb_x0 = std::max(halo, std::atoi(e));
    s->tb_x0 += (s->tb_x0 - halo) % s->tb_v;  // the strip's first column is a multiple of V
    VL_TRY(build_passes(s, rows));
    s->nbuf = s->tb_k + 1;
    for (int j = 2; j < s->nbuf; ++j) {
      VL_HIP(hipMalloc(&s->buf[j], s->pop_bytes));
      s->device_bytes += (int64_t)s->pop_bytes;
    }
  }
  if (s->split) {
    VL_HIP(hipStreamCreateWithFlags(&s->side, hipStreamNonBlocking));
    VL_HIP(hipEventCreateWithFlags(&s->ev_fork, hipEventDisableTiming));
    VL_HIP(hipEventCreateWithFlags(&s->ev_join, hipEventDisableTiming));
    for (int i = 0; i < 2; ++i) {
      VL_HIP(hipEventCreateWithFlags(&s->ev_far[i], hipEventDisableTiming));
      VL_HIP(hipEventCreateWithFlags(&s->ev_near[i], hipEventDisableTiming));
    }
  }
  if (!outlet_nodes_match(tb, t.npb))
    return set_error(VL_EINVAL, "internal: outlet-controller nodes do not match the pressure-boundary nodes");
  VL_HIP(hipMalloc((void**)&t.pvel, (size_t)std::max(t.np, 1) * 2 * sizeof(double)));
  VL_HIP(hipMemset(t.pvel, 0, (size_t)std::max(t.np, 1) * 2 * sizeof(double)));
  VL_HIP(hipMalloc((void**)&t.drho, sizeof(double)));
  VL_HIP(hipMemset(t.drho, 0, sizeof(double)));
  s->last_force.assign((size_t)t.ngroups * 2, 0.0);
  VL_TRY(ensure_hist(s, RUN_CHUNK));  // fixed for the sim's lifetime: graphs hold the pointer
  VL_HIP(hipMalloc((void**)&s->d_ctl, RUN_CHUNK * sizeof(StepCtl)));
  VL_HIP(hipHostMalloc((void**)&s->h_ctl, RUN_CHUNK * sizeof(StepCtl), hipHostMallocDefault));
  VL_HIP(hipMalloc((void**)&s->d_kbase, 2 * sizeof(int64_t)));
  VL_HIP(hipMemset(s->d_kbase, 0, 2 * sizeof(int64_t)));
  s->d_kzero = s->d_kbase + 1;
  s->use_graphs = true;
  if (const char* e = std::getenv("VL_GRAPHS")) s->use_graphs = e[0] != '0';
  // rest state, as Sim::from_builder leaves it (src/sim.rs:352)
  std::vector<char> host(s->pop_bytes, 0);
  for (int64_t y = 0; y < d.ny; ++y)
    for (int64_t x = 0; x < d.nx; ++x) host_store_eq(s, host, x, y, 1.0, 0.0, 0.0);
  VL_TRY(write_raw(s, host));
  s->t = 0;
  *out = guard.release();
  return VL_OK;
}


static vl_status ensure_events(vl_sim* s) {
  if (!s->ev_chunk[0]) {
    VL_HIP(hipEventCreate(&s->ev_chunk[0]));
    VL_HIP(hipEventCreate(&s->ev_chunk[1]));
  }
  while ((int64_t)s->ev.size() < 2 * RUN_CHUNK) {
    hipEvent_t e;
    VL_HIP(hipEventCreate(&e));
    s->ev.push_back(e);
  }
  return VL_OK;
}

// Adds up the event pairs recorded in the chunk that just completed.
static vl_status collect_times(vl_sim* s, int64_t steps) {
  for (int64_t k = 0; k + 1 < s->ev_used; k += 2) {
    float ms = 0.0f;
    VL_HIP(hipEventElapsedTime(&ms, s->ev[k], s->ev[k + 1]));
    s->bulk_ms += ms;
    s->bulk_launches += 1;
  }
  float ms = 0.0f;
  VL_HIP(hipEventElapsedTime(&ms, s->ev_chunk[0], s->ev_chunk[1]));
  s->total_ms += ms;
  s->timed_steps += steps;
  s->ev_used = 0;
  return VL_OK;
}

// Step graphs: a window of GRAPH_K steps (whole groups) captured once per (current buffer, step
// parity) and replayed; the kernel arguments of a window do not change between replays
// (controls and history are indexed through the device counter kbase), so no per-step host
// work remains.  A run shorter than a window replays a graph of its whole groups when one has
// been instantiated for that state and length (vl_sim_prepare), and launches them directly
// otherwise.
static constexpr int GRAPH_K = 32;
static int window_steps(const vl_sim* s) { return GRAPH_K / s->tb_k * s->tb_k; }  // whole groups

// The graph of `steps` steps (whole groups) from the current state, captured and instantiated
// on first use; `after` = the current buffer after it.
static vl_status get_step_graph(vl_sim* s, int steps, hipGraphExec_t* out, int* after) {
  const std::pair<int, int> key(s->cur * 2 + (int)(s->t & 1), steps);
  auto found = s->graphs.find(key);
  if (found != s->graphs.end()) {
    *out = found->second.first;
    *after = found->second.second;
    return VL_OK;
  }
  const int cur0 = s->cur;
  const int64_t t0 = s->t;
  // A graph holds the kernel arguments it was captured with, so it always binds the handle's
  // own control table and force-history staging (allocated with the simulation and kept for its
  // lifetime), whatever the current run set: before the first run those pointers are unset, and
  // vl_sim_run_noalloc points them at its scratch area.
  if (!s->d_ctl || !s->dhist) return set_error(VL_EINVAL, "internal: step graph without the run buffers");
  StepCtl* const run_ctl0 = s->run_ctl;
  double* const run_hist0 = s->run_hist;
  s->run_ctl = s->d_ctl;
  s->run_hist = s->dhist;
  hipGraph_t gph = nullptr;
  hipError_t e = hipStreamBeginCapture(s->stream, hipStreamCaptureModeThreadLocal);
  vl_status st = VL_OK;
  if (e == hipSuccess) {
    const int gs = group_steps(s);
    for (int k = 0; k < steps && st == VL_OK; k += gs) st = some_steps(s, gs, s->d_kbase, k, false);
    if (st == VL_OK) st = join_overlap(s);
    if (st == VL_OK) hipLaunchKernelGGL(k_advance, dim3(1), dim3(1), 0, s->stream, s->d_kbase, (int64_t)steps);
    e = hipStreamEndCapture(s->stream, &gph);
  }
  const int cur_after = s->cur;
  s->cur = cur0;  // capture recorded the steps without running them
  s->t = t0;
  s->run_ctl = run_ctl0;
  s->run_hist = run_hist0;
  if (st != VL_OK) {
    if (gph) (void)hipGraphDestroy(gph);
    return st;
  }
  if (e != hipSuccess) return set_error(VL_EDEVICE, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
  hipGraphExec_t ge = nullptr;
  e = hipGraphInstantiate(&ge, gph, nullptr, nullptr, 0);
  (void)hipGraphDestroy(gph);
  if (e != hipSuccess) return set_error(VL_EDEVICE, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
  // upload it now, so that its first launch costs what a replay does (a graph prepared ahead of a
  // run, vl_sim_prepare, then keeps that cost out of the run)
  e = hipGraphUpload(ge, s->stream);
  if (e != hipSuccess) {
    (void)hipGraphExecDestroy(ge);
    return set_error(VL_EDEVICE, std::string("hipGraphUpload: ") + hipGetErrorString(e));
  }
  s->graphs[key] = {ge, cur_after};
  *out = ge;
  *after = cur_after;
  return VL_OK;
}

static bool has_step_graph(const vl_sim* s, int steps) {
  return s->graphs.count(std::pair<int, int>(s->cur * 2 + (int)(s->t & 1), steps)) != 0;
}

// Device buffers of a run: the control table and the force-history staging of chunks of up to
// `chunk` steps, and whether step-window graphs may be replayed (graphs hold the handle's own
// buffers, so only runs on those may use them).
struct RunBufs {
  StepCtl* ctl;
  double* hist;
  int64_t chunk;
  bool graphs;
};

static vl_status run_steps(vl_sim* s, int64_t nsteps, const vl_ctl* ctl, double* force_hist, const RunBufs& rb) {
  if (nsteps < 0) return set_error(VL_EINVAL, "negative step count");
  const int64_t ng = s->dt.ngroups;
  const int64_t chunk = rb.chunk;
  s->run_ctl = rb.ctl;
  s->run_hist = rb.hist;
  if (s->timing) VL_TRY(ensure_events(s));
  if (nsteps > 0 && s->dt.np > 0) VL_TRY(prime_outlet(s));
  // Graph replay is used without kernel timing (events cannot sit inside the window).
  const bool graphs = rb.graphs && s->use_graphs && !s->timing;
  for (int64_t k0 = 0; k0 < nsteps; k0 += chunk) {
    const int64_t m = std::min(chunk, nsteps - k0);
    // controls of this chunk -> pinned staging -> device table
    for (int64_t k = 0; k < m; ++k) {
      StepCtl& c = s->h_ctl[k];
      if (ctl) {
        for (int g = 0; g < 4; ++g) c.vscale[g] = ctl[k0 + k].vscale[g];
        c.target = ctl[k0 + k].target;
        c.has_target = ctl[k0 + k].has_target;
      } else {
        for (int g = 0; g < 4; ++g) c.vscale[g] = s->vscale[g];
        c.target = s->target;
        c.has_target = s->has_target;
      }
      if (s->dt.np == 0) c.has_target = 0;
    }
    VL_HIP(hipMemcpyAsync(s->run_ctl, s->h_ctl, (size_t)m * sizeof(StepCtl), hipMemcpyHostToDevice, s->stream));
    if (s->timing) {
      s->ev_used = 0;
      VL_HIP(hipEventRecord(s->ev_chunk[0], s->stream));
    }
    int64_t k = 0;
    const int win = window_steps(s), gs = group_steps(s);
    if (graphs && m >= gs) {
      VL_HIP(hipMemsetAsync(s->d_kbase, 0, sizeof(int64_t), s->stream));
      int after = 0;
      for (; k + win <= m; k += win) {
        hipGraphExec_t ge;
        VL_TRY(get_step_graph(s, win, &ge, &after));
        VL_HIP(hipGraphLaunch(ge, s->stream));
        s->cur = after;
        s->t += win;
      }
      // the whole groups left, as one prepared graph
      const int rest = (int)((m - k) / gs * gs);
      if (rest > 0 && has_step_graph(s, rest)) {
        hipGraphExec_t ge;
        VL_TRY(get_step_graph(s, rest, &ge, &after));
        VL_HIP(hipGraphLaunch(ge, s->stream));
        s->cur = after;
        s->t += rest;
        k += rest;
      }
    }
    while (k < m) {  // remainder, launched directly
      const int n = m - k >= group_steps(s) ? group_steps(s) : 1;
      VL_TRY(some_steps(s, n, s->d_kzero, (int)k, true));
      k += n;
    }
    VL_TRY(join_overlap(s));
    if (ng > 0) VL_TRY(finish_forces(s, s->run_hist + (m - 1) * ng * 2));
    if (s->timing) VL_HIP(hipEventRecord(s->ev_chunk[1], s->stream));
    if (ng > 0) {
      if (force_hist)
        VL_HIP(hipMemcpyAsync(force_hist + k0 * ng * 2, s->run_hist, (size_t)(m * ng * 2) * sizeof(double),
                              hipMemcpyDeviceToHost, s->stream));
      VL_HIP(hipMemcpyAsync(s->last_force.data(), s->run_hist + (m - 1) * ng * 2, (size_t)(ng * 2) * sizeof(double),
                            hipMemcpyDeviceToHost, s->stream));
    }
    VL_HIP(hipStreamSynchronize(s->stream));
    if (s->timing) VL_TRY(collect_times(s, m));
    const StepCtl& last = s->h_ctl[m - 1];
    s->has_target = last.has_target;
    s->target = last.target;
    for (int g = 0; g < 4; ++g) s->vscale[g] = last.vscale[g];
  }
  if (nsteps > 0 && s->has_target && s->dt.np > 0)
    VL_HIP(hipMemcpy(&s->outlet_density_offset, s->dt.drho, sizeof(double), hipMemcpyDeviceToHost));
  else if (nsteps > 0)
    s->outlet_density_offset = 0.0;
  return VL_OK;
}

// Bandwidth of a plain device copy of `bytes` (read + write counted), best of `iters`.
static vl_status copy_bandwidth(size_t bytes, int iters, double* gbs) {
  const size_t n = bytes / 16;
  if (n == 0 || iters < 1) return set_error(VL_EINVAL, "copy size and iterations must be positive");
  float4 *a = nullptr, *b = nullptr;
  hipStream_t st = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  vl_status rc = VL_OK;
  auto run = [&]() -> vl_status {
    VL_HIP(hipMalloc((void**)&a, n * 16));
    VL_HIP(hipMalloc((void**)&b, n * 16));
    VL_HIP(hipMemset(a, 0, n * 16));
    VL_HIP(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
    VL_HIP(hipEventCreate(&e0));
    VL_HIP(hipEventCreate(&e1));
    if ((n + 255) / 256 > 0x7fffffff) return set_error(VL_EINVAL, "copy size too large");
    const unsigned grid = (unsigned)((n + 255) / 256);
    float best = 1e30f;
    for (int it = 0; it <= iters; ++it) {  // first launch warms up
      VL_HIP(hipEventRecord(e0, st));
      hipLaunchKernelGGL(k_copy, dim3(grid), dim3(256), 0, st, a, b, n);
      VL_HIP(hipGetLastError());
      VL_HIP(hipEventRecord(e1, st));
      VL_HIP(hipEventSynchronize(e1));
      float ms = 0.0f;
      VL_HIP(hipEventElapsedTime(&ms, e0, e1));
      if (it > 0) best = std::min(best, ms);
    }
    *gbs = 2.0 * (double)(n * 16) / (best * 1e-3) / 1e9;
    return VL_OK;
  };
  rc = run();
  if (a) (void)hipFree(a);
  if (b) (void)hipFree(b);
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (st) (void)hipStreamDestroy(st);
  return rc;
}

}  // namespace vl

using namespace vl;

// ==========================================================================================
// C ABI
// ==========================================================================================
extern "C" {

int vl_device_count(void) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) return 0;
  return n;
}

vl_status vl_sim_create(vl_builder* b, int precision, vl_sim** out) {
  if (!b || !out) {
    if (b) vl_builder_free(b);
    return set_error(VL_EINVAL, "null argument");
  }
  *out = nullptr;
  vl_status st = create(b->d, precision, out);
  vl_builder_free(b);
  return st;
}

void vl_sim_destroy(vl_sim* s) { free_sim(s); }

vl_status vl_plan_passes(const uint8_t* flags, int nx, int ny, int k, int v, int rows, int x0, int waves,
                         int join, int32_t* tasks, int64_t max_tasks, uint32_t* listed, int64_t max_listed,
                         int64_t* counts) {
  if (!flags || !counts || nx < 3 || ny < 3 || k < 2 || k > TB_MAX_K || (v != 1 && v != 2) || x0 < tb_halo(k, v) ||
      (x0 - tb_halo(k, v)) % v != 0 || waves < 0 || waves > 16 || (join && waves < 2))
    return set_error(VL_EINVAL,
                     "vl_plan_passes: flags, counts, nx, ny >= 3, k in 2..5, v in 1..2, x0 = halo + a multiple of v, "
                     "waves in 0..16 (at least 2 with join)");
  PassPlan plan;
  std::vector<uint8_t> fl(flags, flags + (size_t)nx * ny);
  VL_TRY(plan_passes(fl, nx, ny, round_up(nx, 64), k, v, rows, x0, join ? waves : 0, plan));
  if (waves > 0) {  // the lockstep bits as vl_sim_create sets them
    const TaskRanges ranges = pass_ranges((int64_t)plan.tasks.size(), plan.nfar);
    mark_lockstep(plan.tasks, nullptr, ranges, waves);
    VL_TRY(check_lockstep(plan.tasks, nullptr, ranges, waves, TB_LANES * v));
  }
  counts[0] = (int64_t)plan.tasks.size();
  counts[1] = plan.nfar;
  counts[2] = (int64_t)plan.listed.size();
  counts[3] = plan.covered;
  counts[4] = plan.njoined_far;
  counts[5] = plan.njoined_near;
  if (tasks)
    for (int64_t i = 0; i < std::min<int64_t>(max_tasks, counts[0]); ++i) {
      const int4& t = plan.tasks[(size_t)i];
      tasks[4 * i] = t.x;
      tasks[4 * i + 1] = t.y;
      tasks[4 * i + 2] = t.z;
      tasks[4 * i + 3] = t.w;
    }
  if (listed)
    for (int64_t i = 0; i < std::min<int64_t>(max_listed, counts[2]); ++i) listed[i] = plan.listed[(size_t)i];
  return VL_OK;
}

vl_status vl_pass_lockstep(int32_t* tasks, const int32_t* owner, int64_t ntasks, const int64_t* ranges, int nranges,
                           int waves, int check_only) {
  if ((!tasks && ntasks > 0) || ntasks < 0 || (!ranges && nranges > 0) || nranges < 0 || waves < 1)
    return set_error(VL_EINVAL, "vl_pass_lockstep: tasks, ntasks >= 0, ranges, waves >= 1");
  std::vector<int4> t((size_t)ntasks);
  for (int64_t i = 0; i < ntasks; ++i)
    t[(size_t)i] = make_int4(tasks[4 * i], tasks[4 * i + 1], tasks[4 * i + 2], tasks[4 * i + 3]);
  TaskRanges r;
  for (int i = 0; i < nranges; ++i) r.emplace_back(ranges[2 * i], ranges[2 * i + 1]);
  for (const auto& [b, e] : r)
    if (b < 0 || e > ntasks || b > e) return set_error(VL_EINVAL, "vl_pass_lockstep: a range outside the list");
  if (!check_only) {
    mark_lockstep(t, owner, r, waves);
    for (int64_t i = 0; i < ntasks; ++i) tasks[4 * i + 3] = t[(size_t)i].w;
  }
  return check_lockstep(t, owner, r, waves, -1);
}

vl_status vl_sim_info_get(vl_sim* s, vl_sim_info* o) {
  if (!s || !o) return set_error(VL_EINVAL, "null argument");
  std::memset(o, 0, sizeof(*o));
  o->nx = s->g.nx;
  o->ny = s->g.ny;
  o->fluid_nodes = s->tab.fluid_nodes;
  o->link_count = s->dt.nl;
  o->bnode_count = s->dt.nb;
  o->pnode_count = s->dt.np;
  o->force_groups = s->dt.ngroups;
  o->t = s->t;
  o->precision = s->precision;
  o->collision = s->collision;
  o->omega_plus = s->coll.op;
  o->omega_minus = s->coll.om;
  o->outlet_density_offset = s->outlet_density_offset;
  o->device_bytes = s->device_bytes;
  o->split = s->split ? 1 : 0;
  o->pass_steps = s->tb_k;
  o->pass_nodes = s->tb_nodes;
  o->pass_tasks = s->ntasks;
  o->listed_nodes = s->nlisted;
  o->pass_rows = s->tb_rows;
  o->pass_joined_tasks = s->tb_joined_far + s->tb_joined_near;
  o->pass_far_nodes = s->tb_far_nodes;
  o->pass_far_tasks = s->tb_overlap ? s->ntasks_far : 0;
  return VL_OK;
}

vl_status vl_sim_init_macro(vl_sim* s, const double* rho, const double* ux, const double* uy) {
  if (!s || !rho || !ux || !uy) return set_error(VL_EINVAL, "null argument");
  std::vector<char> host(s->pop_bytes, 0);
  const int64_t nx = s->g.nx;
  for (int64_t y = 0; y < s->g.ny; ++y)
    for (int64_t x = 0; x < nx; ++x) {
      const int64_t k = y * nx + x;
      if (s->tab.flag[k] == 2)
        host_store_eq(s, host, x, y, 1.0, 0.0, 0.0);
      else
        host_store_eq(s, host, x, y, rho[k], ux[k], uy[k]);
    }
  VL_TRY(write_raw(s, host));
  s->t = 0;
  return VL_OK;
}

vl_status vl_sim_set_vscale(vl_sim* s, const double* v) {
  if (!s || !v) return set_error(VL_EINVAL, "null argument");
  for (int g = 0; g < 4; ++g) s->vscale[g] = v[g];
  return VL_OK;
}

vl_status vl_sim_set_outflow_target(vl_sim* s, int enabled, double target) {
  if (!s) return set_error(VL_EINVAL, "null argument");
  s->has_target = enabled != 0;
  s->target = target;
  return VL_OK;
}

vl_status vl_sim_step(vl_sim* s, int64_t nsteps) {
  if (!s) return set_error(VL_EINVAL, "null argument");
  return run_steps(s, nsteps, nullptr, nullptr, RunBufs{s->d_ctl, s->dhist, RUN_CHUNK, true});
}

vl_status vl_sim_run(vl_sim* s, int64_t nsteps, const vl_ctl* ctl, double* force_hist) {
  if (!s || (!ctl && nsteps > 0)) return set_error(VL_EINVAL, "null argument");
  return run_steps(s, nsteps, ctl, force_hist, RunBufs{s->d_ctl, s->dhist, RUN_CHUNK, true});
}

vl_status vl_sim_prepare(vl_sim* s, int64_t nsteps) { return vl_sim_prepare_after(s, 0, nsteps); }

vl_status vl_sim_prepare_after(vl_sim* s, int64_t skip, int64_t nsteps) {
  if (!s || nsteps < 0 || skip < 0)
    return set_error(VL_EINVAL, "vl_sim_prepare: null simulation or negative step count");
  if (!s->use_graphs) return VL_OK;
  // the state sequence of run_steps over nsteps, without running, from the state `skip` steps
  // ahead (every step of a run moves the current buffer on by one): whole windows, one graph of
  // the whole groups left in each chunk, then single steps
  const int cur0 = s->cur;
  const int64_t t0 = s->t;
  s->cur = (int)((s->cur + skip) % s->nbuf);
  s->t += skip;
  const int win = window_steps(s), gs = group_steps(s);
  vl_status st = VL_OK;
  for (int64_t k0 = 0; k0 < nsteps && st == VL_OK; k0 += RUN_CHUNK) {
    const int64_t m = std::min<int64_t>(RUN_CHUNK, nsteps - k0);
    int64_t k = 0;
    hipGraphExec_t ge;
    int after = 0;
    for (; k + win <= m && st == VL_OK; k += win) {
      st = get_step_graph(s, win, &ge, &after);
      s->cur = after;
      s->t += win;
    }
    const int rest = (int)((m - k) / gs * gs);
    if (rest > 0 && st == VL_OK) {
      st = get_step_graph(s, rest, &ge, &after);
      s->cur = after;
      s->t += rest;
      k += rest;
    }
    for (; k < m; ++k) {  // single steps (launch_step's state change)
      s->cur = (s->cur + 1) % s->nbuf;
      s->t += 1;
    }
  }
  s->cur = cur0;
  s->t = t0;
  // the graphs' uploads run on the stream: wait for them, so that a run timed after this call
  // starts on a quiet stream
  if (st == VL_OK) VL_HIP(hipStreamSynchronize(s->stream));
  return st;
}

vl_status vl_sim_run_noalloc(vl_sim* s, int64_t nsteps, const vl_ctl* ctl, double* force_hist) {
  if (!s || (!ctl && nsteps > 0)) return set_error(VL_EINVAL, "null argument");
  if (nsteps > NOALLOC_MAX_STEPS)
    return set_error(VL_EINVAL, "vl_sim_run_noalloc: at most " + std::to_string(NOALLOC_MAX_STEPS) +
                                    " steps per call (got " + std::to_string(nsteps) + ")");
  if (s->dt.ngroups > NOALLOC_MAX_GROUPS)
    return set_error(VL_EINVAL, "vl_sim_run_noalloc: at most " + std::to_string(NOALLOC_MAX_GROUPS) +
                                    " force groups (the simulation has " + std::to_string(s->dt.ngroups) + ")");
  if (s->timing) return set_error(VL_EINVAL, "vl_sim_run_noalloc: switch kernel timing off first");
  char* p = nullptr;
  VL_TRY(noalloc_scratch(s->device, &p));
  return run_steps(s, nsteps, ctl, force_hist,
                   RunBufs{(StepCtl*)p, (double*)(p + NOALLOC_CTL_BYTES), NOALLOC_MAX_STEPS, false});
}

vl_status vl_device_copy_bandwidth(size_t bytes, int iters, double* gbs) {
  if (!gbs) return set_error(VL_EINVAL, "null argument");
  return copy_bandwidth(bytes, iters, gbs);
}

vl_status vl_sim_set_timing(vl_sim* s, int on) {
  if (!s) return set_error(VL_EINVAL, "null argument");
  s->timing = on != 0;
  s->bulk_ms = s->total_ms = 0.0;
  s->bulk_launches = s->timed_steps = 0;
  s->ev_used = 0;
  return VL_OK;
}

vl_status vl_sim_get_timing(vl_sim* s, double* bulk_ms, int64_t* bulk_launches, double* total_ms,
                            int64_t* steps) {
  if (!s || !bulk_ms || !bulk_launches || !total_ms || !steps) return set_error(VL_EINVAL, "null argument");
  *bulk_ms = s->bulk_ms;
  *bulk_launches = s->bulk_launches;
  *total_ms = s->total_ms;
  *steps = s->timed_steps;
  return VL_OK;
}

vl_status vl_sim_get_force(vl_sim* s, double* out) {
  if (!s || !out) return set_error(VL_EINVAL, "null argument");
  std::copy(s->last_force.begin(), s->last_force.end(), out);
  return VL_OK;
}

vl_status vl_sim_read(vl_sim* s, double* f) {
  if (!s || !f) return set_error(VL_EINVAL, "null argument");
  std::vector<char> host;
  VL_TRY(read_raw(s, host));
  const int64_t nx = s->g.nx, ny = s->g.ny, n = nx * ny;
  for (int i = 0; i < Q; ++i)
    for (int64_t y = 0; y < ny; ++y)
      for (int64_t x = 0; x < nx; ++x) f[i * n + y * nx + x] = host_load(s, host, i, y * s->g.pitch + x);
  return VL_OK;
}

vl_status vl_sim_write(vl_sim* s, const double* f) {
  if (!s || !f) return set_error(VL_EINVAL, "null argument");
  std::vector<char> host(s->pop_bytes, 0);
  const int64_t nx = s->g.nx, ny = s->g.ny, n = nx * ny;
  for (int i = 0; i < Q; ++i)
    for (int64_t y = 0; y < ny; ++y)
      for (int64_t x = 0; x < nx; ++x) {
        const int64_t e = i * s->g.plane + y * s->g.pitch + x;
        const double v = f[i * n + y * nx + x];
        if (s->precision == VL_F64)
          ((double*)host.data())[e] = v;
        else if (s->precision == VL_F32)
          ((float*)host.data())[e] = (float)(v - wgt(i));
        else
          ((_Float16*)host.data())[e] = (_Float16)((float)(v - wgt(i)) * F16_SCALE);
      }
  return write_raw(s, host);
}

// Sim::macros (src/sim.rs:527-534) for every node.
vl_status vl_sim_macros(vl_sim* s, double* rho, double* ux, double* uy) {
  if (!s || !rho || !ux || !uy) return set_error(VL_EINVAL, "null argument");
  std::vector<char> host;
  VL_TRY(read_raw(s, host));
  const int64_t nx = s->g.nx, ny = s->g.ny;
  for (int64_t y = 0; y < ny; ++y)
    for (int64_t x = 0; x < nx; ++x) {
      const int64_t k = y * nx + x;
      if (s->tab.flag[k] == 2) {
        rho[k] = 1.0;
        ux[k] = 0.0;
        uy[k] = 0.0;
        continue;
      }
      double f[Q];
      for (int i = 0; i < Q; ++i) f[i] = host_load(s, host, i, y * s->g.pitch + x);
      double r, jx, jy;
      Exact::moments(f, r, jx, jy);
      rho[k] = r;
      ux[k] = jx / r - 0.5 * s->accel[0];
      uy[k] = jy / r - 0.5 * s->accel[1];
    }
  return VL_OK;
}

// FNV-1a over the little-endian bytes of every fluid population, node-major, direction-minor
// (src/sim.rs:654-668).
vl_status vl_sim_state_hash(vl_sim* s, uint64_t* out) {
  if (!s || !out) return set_error(VL_EINVAL, "null argument");
  std::vector<char> host;
  VL_TRY(read_raw(s, host));
  uint64_t h = 0xcbf29ce484222325ull;
  const int64_t nx = s->g.nx, ny = s->g.ny;
  for (int64_t y = 0; y < ny; ++y)
    for (int64_t x = 0; x < nx; ++x) {
      if (s->tab.flag[y * nx + x] == 2) continue;
      for (int i = 0; i < Q; ++i) {
        const double v = host_load(s, host, i, y * s->g.pitch + x);
        uint64_t bits;
        std::memcpy(&bits, &v, 8);
        for (int k = 0; k < 8; ++k) {
          h ^= (bits >> (8 * k)) & 0xffu;
          h *= 0x100000001b3ull;
        }
      }
    }
  *out = h;
  return VL_OK;
}

namespace {
struct Compensated {  // Kahan-Babuska-Neumaier (the reference's diagnostic sum, src/sim.rs:676-691)
  double sum = 0.0, c = 0.0;
  void add(double v) {
    const double t = sum + v;
    c += std::fabs(sum) >= std::fabs(v) ? (sum - t) + v : (v - t) + sum;
    sum = t;
  }
  double value() const { return sum + c; }
};
}  // namespace

vl_status vl_sim_total_mass(vl_sim* s, double* out) {
  if (!s || !out) return set_error(VL_EINVAL, "null argument");
  std::vector<char> host;
  VL_TRY(read_raw(s, host));
  Compensated m;
  const int64_t nx = s->g.nx, ny = s->g.ny;
  for (int64_t y = 0; y < ny; ++y)
    for (int64_t x = 0; x < nx; ++x)
      if (s->tab.flag[y * nx + x] != 2)
        for (int i = 0; i < Q; ++i) m.add(host_load(s, host, i, y * s->g.pitch + x));
  *out = m.value();
  return VL_OK;
}

vl_status vl_sim_total_momentum(vl_sim* s, double* out) {
  if (!s || !out) return set_error(VL_EINVAL, "null argument");
  std::vector<char> host;
  VL_TRY(read_raw(s, host));
  Compensated px, py;
  const int64_t nx = s->g.nx, ny = s->g.ny;
  for (int64_t y = 0; y < ny; ++y)
    for (int64_t x = 0; x < nx; ++x)
      if (s->tab.flag[y * nx + x] != 2)
        for (int i = 1; i < Q; ++i) {
          const double f = host_load(s, host, i, y * s->g.pitch + x);
          px.add((double)cx(i) * f);
          py.add((double)cy(i) * f);
        }
  out[0] = px.value();
  out[1] = py.value();
  return VL_OK;
}

// Sim::vorticity (src/sim.rs:581-620): dv/dx - du/dy by central differences, one-sided where one
// neighbour along the axis is not fluid (or lies past the domain edge; the differences do not
// wrap), 0 where neither is and in non-fluid cells; out: nx * ny values, node (x, y) at y nx + x.
vl_status vl_sim_vorticity(vl_sim* s, float* out) {
  if (!s || !out) return set_error(VL_EINVAL, "null argument");
  std::vector<char> host;
  VL_TRY(read_raw(s, host));
  const int64_t nx = s->g.nx, ny = s->g.ny;
  std::vector<double> u((size_t)(2 * nx * ny), 0.0);
  for (int64_t y = 0; y < ny; ++y)
    for (int64_t x = 0; x < nx; ++x) {
      const int64_t k = y * nx + x;
      if (s->tab.flag[k] == 2) continue;
      double f[Q];
      for (int i = 0; i < Q; ++i) f[i] = host_load(s, host, i, y * s->g.pitch + x);
      double r, jx, jy;
      Exact::moments(f, r, jx, jy);
      u[2 * k] = jx / r - 0.5 * s->accel[0];
      u[2 * k + 1] = jy / r - 0.5 * s->accel[1];
    }
  auto fluid = [&](int64_t x, int64_t y) {
    return x >= 0 && y >= 0 && x < nx && y < ny && s->tab.flag[y * nx + x] != 2;
  };
  // d(component comp)/d(axis (dx, dy)) at (x, y) from whichever neighbours are fluid
  auto diff = [&](int64_t x, int64_t y, int dx, int dy, int comp) {
    const auto at = [&](int64_t xx, int64_t yy) { return u[(size_t)(2 * (yy * nx + xx) + comp)]; };
    const bool p = fluid(x + dx, y + dy), m = fluid(x - dx, y - dy);
    if (p && m) return 0.5 * (at(x + dx, y + dy) - at(x - dx, y - dy));
    if (p) return at(x + dx, y + dy) - at(x, y);
    if (m) return at(x, y) - at(x - dx, y - dy);
    return 0.0;
  };
  for (int64_t y = 0; y < ny; ++y)
    for (int64_t x = 0; x < nx; ++x)
      out[y * nx + x] = fluid(x, y) ? (float)(diff(x, y, 1, 0, 1) - diff(x, y, 0, 1, 0)) : 0.0f;
  return VL_OK;
}

// Sim::is_fluid (src/sim.rs:381-383); a node outside the domain is an error (the reference
// panics on the index).
vl_status vl_sim_is_fluid(vl_sim* s, int64_t x, int64_t y, int* out) {
  if (!s || !out) return set_error(VL_EINVAL, "null argument");
  if (x < 0 || y < 0 || x >= s->g.nx || y >= s->g.ny) return set_error(VL_EINVAL, "node outside the domain");
  *out = s->tab.flag[y * s->g.nx + x] != 2;
  return VL_OK;
}

// Sim::periodic (src/sim.rs:670-672).
vl_status vl_sim_periodic(vl_sim* s, int* px, int* py) {
  if (!s || !px || !py) return set_error(VL_EINVAL, "null argument");
  *px = s->g.periodic_x != 0;
  *py = s->g.periodic_y != 0;
  return VL_OK;
}

vl_status vl_sim_max_speed(vl_sim* s, double* out) {
  if (!s || !out) return set_error(VL_EINVAL, "null argument");
  std::vector<char> host;
  VL_TRY(read_raw(s, host));
  double m = 0.0;
  const int64_t nx = s->g.nx, ny = s->g.ny;
  for (int64_t y = 0; y < ny; ++y)
    for (int64_t x = 0; x < nx; ++x) {
      if (s->tab.flag[y * nx + x] == 2) continue;
      double f[Q];
      for (int i = 0; i < Q; ++i) f[i] = host_load(s, host, i, y * s->g.pitch + x);
      double r, jx, jy;
      Exact::moments(f, r, jx, jy);
      const double ux = jx / r - 0.5 * s->accel[0], uy = jy / r - 0.5 * s->accel[1];
      const double sp = std::sqrt(ux * ux + uy * uy);
      if (std::isnan(sp)) {
        *out = NAN;
        return VL_OK;
      }
      m = std::max(m, sp);
    }
  *out = m;
  return VL_OK;
}

}  // extern "C"
