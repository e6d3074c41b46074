// Device-side simulation state shared by the HIP translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <map>
#include <string>
#include <utility>
#include <vector>

#include "setup.h"

namespace vl {

#define VL_HIP(call)                                                                         \
  do {                                                                                       \
    hipError_t e_ = (call);                                                                  \
    if (e_ != hipSuccess)                                                                    \
      return ::vl::set_error(VL_EDEVICE, std::string(#call) + ": " + hipGetErrorString(e_)); \
  } while (0)

// Per-step controls as the kernels see them (a copy of vl_ctl).
struct StepCtl {
  double vscale[4];
  double target;
  int has_target;
};

// Collision parameters.
struct Coll {
  double op, om, gx, gy;
  float opf, omf, gxf, gyf;
  int model;  // Model
};

// Geometry of the device arrays: population i of node (x, y) lives at
// base[i * plane + y * pitch + x]; pitch >= nx is padded for aligned rows and `plane`
// leaves a gap between planes so the nine streams do not alias in the HBM channels.
struct Geom {
  int32_t nx, ny, pitch;
  int32_t periodic_x, periodic_y;
  int64_t plane;
};

struct DevTables {
  int32_t nb = 0, nl = 0, np = 0, nslots = 0, ngroups = 0;
  int32_t npb = 0;  // boundary nodes (listed first) whose links use the outlet controller
  int32_t *bx = nullptr, *by = nullptr, *bfirst = nullptr;
  uint8_t *dir = nullptr, *kind = nullptr, *vgroup = nullptr, *lflags = nullptr;
  int8_t *lnx = nullptr, *lny = nullptr;
  int32_t* fslot = nullptr;
  double *q = nullptr, *uwx = nullptr, *uwy = nullptr, *cu6 = nullptr, *rho = nullptr;
  int32_t *px = nullptr, *py = nullptr;
  int8_t *pnx = nullptr, *pny = nullptr;
  int32_t* gfirst = nullptr;   // ngroups+1 slot offsets
  double* fcontrib = nullptr;  // 2 steps x nslots x (x, y)
  double* drho = nullptr;      // 1 value: outlet density offset of the current step
  double* pvel = nullptr;      // 2 steps x np outward velocities of the outlet-controller nodes
};

// K-step pass plan of one domain (plan_passes in sim.hip): the strips' tasks (first output
// column x, first output row y, rows z, output columns | ring-edge bits << 8 in w), far tasks
// first; the interior nodes that no task covers (y << 16 | x); the covered node counts.
struct PassPlan {
  std::vector<int4> tasks;
  std::vector<uint32_t> listed;
  int nfar = 0;
  int njoined_far = 0, njoined_near = 0;  // tasks of joined groups at the start of each range
  int64_t covered = 0, far_nodes = 0;
};
// flags: nx * ny bytes, 0 interior, 1 boundary, 2 solid; K steps per pass, V columns per lane,
// tasks of at most `rows` output rows, strips aligned from column x0_grid; `join` = W > 1
// plans joined strips for workgroups of W waves (0: single-wave strips only).
vl_status plan_passes(const std::vector<uint8_t>& flags, int nx, int ny, int64_t pitch, int K, int V, int rows,
                      int x0_grid, int join, PassPlan& plan);

// Lockstep bit (1 << 12 in a task's w) of a pass task list.  The list is launched as one or
// more ranges [begin, end), each as workgroups of W consecutive tasks from `begin` (the last one
// possibly partial).  The bit lets a workgroup's waves sweep in step, one barrier per row, so it
// must only be set where every wave of the workgroup executes the same barriers:
// mark_lockstep sets it on exactly the full workgroups of some range whose W tasks have equal
// first row and row count (and, with `owner`, belong to one case), and clears it elsewhere;
// it leaves joined tasks (bit 13, set by the planner with bit 12) as they are.  check_lockstep
// verifies that property for every range and returns VL_EINVAL otherwise; for joined tasks it
// also requires that they fill whole workgroups whose waves load adjacent runs of `wave_cols`
// columns (64 V; 0: no joined task allowed; -1: 64 or 128).
using TaskRanges = std::vector<std::pair<int64_t, int64_t>>;
void mark_lockstep(std::vector<int4>& tasks, const int32_t* owner, const TaskRanges& ranges, int W);
vl_status check_lockstep(const std::vector<int4>& tasks, const int32_t* owner, const TaskRanges& ranges, int W,
                         int wave_cols = 0);

}  // namespace vl

struct vl_sim {
  int precision = VL_F64;
  vl::Tables tab;          // host copy (flags for readbacks, counts)
  vl::Geom g{};
  vl::Coll coll{};
  int collision = VL_TRT;
  double accel[2] = {0.0, 0.0};
  double vscale[4] = {1.0, 1.0, 1.0, 1.0};
  bool has_target = false;
  double target = 0.0;
  double outlet_density_offset = 0.0;
  int64_t t = 0;
  int device = 0;
  int nlb = 0;                       // boundary workgroups of a step (force workgroup included)
  int fblk = -1;                     // the force workgroup, last of the boundary workgroups
  bool split = false;                // boundary and bulk as two concurrent kernels
  int bulk_v = 1;                    // nodes per thread of the split-mode bulk kernel
  hipStream_t side = nullptr;        // stream of the boundary kernel in split mode
  hipEvent_t ev_fork = nullptr, ev_join = nullptr;
  hipStream_t stream = nullptr;
  void* buf[6] = {};  // population arrays; buf[cur] is A
  int nbuf = 2;                                          // K + 1 with K-step passes
  int cur = 0;
  // K-step passes (temporal blocking of the bulk): a group of K steps runs one kernel that takes
  // the covered interior nodes from t to t+K beside K single steps of every other fluid node.
  int tb_k = 1;                      // steps per pass (1: every step is a single step)
  int tb_rows = 0;                   // output rows per task (upper bound)
  int tb_waves = 8;                  // waves (tasks) per workgroup of the pass kernel
  bool tb_sync = true;               // a workgroup's waves sweep in step where their rows agree
  int tb_x0 = 32;                    // first output column of the first strip
  int tb_v = 1;                      // columns per lane of the pass kernel
  int ntasks = 0;                    // one task per wave: far tasks first, then near ones
  int ntasks_far = 0;
  int64_t tb_far_nodes = 0;          // covered nodes of the far tasks
  bool tb_join = false;              // joined strips (a workgroup's waves sweep one strip)
  int tb_joined_far = 0, tb_joined_near = 0;  // joined tasks at the start of the far / near range
  // Overlap of a group's single steps with the next group's far pass (side stream): whether the
  // side stream has been forked since the last join, and the parity of the group's events.
  bool tb_overlap = true;
  bool ov_open = false;
  int ov_parity = 0;
  hipEvent_t ev_far[2] = {nullptr, nullptr}, ev_near[2] = {nullptr, nullptr};
  int nlisted = 0;                   // interior nodes no task covers
  int64_t tb_nodes = 0;              // interior nodes the tasks cover
  int4* dtasks = nullptr;
  uint32_t* dlisted = nullptr;       // y << 16 | x of the uncovered interior nodes
  size_t elem = 8;
  size_t pop_bytes = 0;  // bytes of one population array
  uint8_t* dflag = nullptr;
  int2* drows = nullptr;  // interior span of every row
  vl::DevTables dt;
  double* dhist = nullptr;  // device force history of a run chunk (hist_cap steps)
  vl::StepCtl* d_ctl = nullptr;  // controls of a run chunk (device)
  // The control table and force-history staging the steps of the current run read and write:
  // d_ctl / dhist, or the library's fixed scratch area for vl_sim_run_noalloc.
  vl::StepCtl* run_ctl = nullptr;
  double* run_hist = nullptr;
  vl::StepCtl* h_ctl = nullptr;  // pinned staging for them
  int64_t* d_kbase = nullptr;    // window counter for replayed step graphs
  int64_t* d_kzero = nullptr;    // a device zero (directly launched steps)
  bool use_graphs = true;
  // Replayed step graphs by (current buffer * 2 + step parity, steps): whole windows, and the
  // shorter runs of whole groups that vl_sim_prepare instantiates; each with the current buffer
  // after it.
  std::map<std::pair<int, int>, std::pair<hipGraphExec_t, int>> graphs;
  int64_t hist_cap = 0;
  std::vector<double> last_force;  // Sim::force after the last step
  int64_t device_bytes = 0;
  // Optional HIP-event timing of the bulk kernel and of whole steps (vl_sim_set_timing).
  bool timing = false;
  std::vector<hipEvent_t> ev;  // per step in a chunk: [bulk start, bulk end]
  hipEvent_t ev_chunk[2] = {nullptr, nullptr};
  int64_t ev_used = 0;
  double bulk_ms = 0.0, total_ms = 0.0;
  int64_t bulk_launches = 0, timed_steps = 0;
};
