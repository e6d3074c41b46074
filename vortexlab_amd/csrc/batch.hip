// Batched solver: hundreds of independent body-in-channel cases of uneven size stepped
// together over a packed tile table (no reference counterpart: the reference runs one case at
// a time, src/cases/tunnel.rs:152-192).  The covered interior of every case advances K steps per
// pass (the single-domain strips, per case); boundary nodes, the interior nodes no task covers
// and the cases that end inside a group of K steps advance in single steps.
//
// Layout: every case keeps its own pitched planes (as a single-domain vl_sim would) inside two
// shared population pools; its boundary tables sit in packed pools and a BatchCase record holds
// pointers into them.  The tile table lists boundary tiles (256 boundary nodes of one case) and
// bulk tiles (TW x TH interior nodes of one case), each sorted by the case's step count so the
// tiles still running at step t are a prefix; two device arrays give that prefix per step.
//
// Per node the arithmetic is the single-domain solver's (step_device.h), the per-step controls
// are the Channel driver's (ramp computed on device with the same operations, kick taken from a
// host-computed table), so a case in a batch gives the same bits as the same case run alone.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <string>
#include <thread>
#include <tuple>
#include <type_traits>
#include <vector>

#include "batch_setup.h"
#include "step_device.h"

namespace vl {

struct BatchCase {
  Geom g;
  Coll c;
  DevTables t;          // pointers into the packed table pools
  int64_t pop_off;      // element offset of the case's array in each population pool
  int64_t flag_off;     // byte offset of the case's flags
  double u_ref, coef_k;  // coefficients = coef_k * force  (2 / (u_ref^2 d))
  uint64_t ramp_steps;
  int64_t kick_start, kick_dur, kick_off;  // kick values for [start, start+dur) at kick_off
  int64_t steps;        // steps this case runs
  int64_t hist_off;     // (cd, cl) history offset, in doubles
  int32_t bnd_tile0;    // its first boundary tile
  int32_t absorbing;
  int32_t inflow;       // initial state when ramp_steps == 0
  int32_t band_w;       // columns of each band beside the inlet and outlet in side arrays (0: none)
  double inflow_u, inflow_h;
  int64_t side_off;     // element offset of the case's band block in each side pool
  int64_t bidx_off;     // offset of the case's band map
  int32_t band_sny;     // rows per band column in the side arrays and the band map
  int32_t pad2_;
  LinkSlots ls;         // slot-major copy of the case's links (meta null: the link tables' columns)
};

constexpr int BT = 256;  // threads of both batch kernels
constexpr int LISTED_NPT = 2;  // listed nodes per thread of k_batch_listed (a tile holds LISTED_NPT * BT)

// The bands of a case in K-step passes: the band_w columns beside the inlet (x = 1 .. band_w) and
// beside the outlet (x = nx - 1 - band_w .. nx - 2), which no pass task covers.  Their
// intermediate levels (1 .. K-1) live in a column-major side array of the case ([plane][band
// column][row], rows padded to band_sny), so that a single step reads and writes whole lines
// of them; level 0 is read from, and level K written to, the population pools like every other
// node.  BandView reads or writes through the side array for band nodes and through the pool
// for every other node; side == null: the pool everywhere.
template <class T>
struct BandView {
  T* main;
  int64_t plane;
  int pitch;
  T* side;
  int bw, xr, sny;
  __device__ inline T* at(int i, int x, int y) const {
    if (side) {
      // The inlet and outlet cells (x = 0 and x = nx - 1) are not fluid: every population pulled
      // from them is replaced by a link rule, so those reads go to the band's first / last
      // column in the side array, beside the node's other loads, instead of a line of their own
      // per row and plane (the value read is never used).
      const int xs = x == 0 ? 1 : (x == xr + bw ? xr + bw - 1 : x);
      const bool lo = xs >= 1 && xs <= bw, hi = xs >= xr && xs < xr + bw;
      if (lo || hi) return side + ((int64_t)(i * 2 * bw + (lo ? xs - 1 : bw + xs - xr)) * sny + y);
    }
    return main + i * plane + ((int64_t)y * pitch + x);
  }
};

__device__ inline double ramp_dev(uint64_t t, uint64_t steps) {  // src/cases/mod.rs:11-18
  if (steps == 0 || t >= steps) return 1.0;
  const double s = (double)t / (double)steps;
  return s * s * (3.0 - 2.0 * s);
}

// Controls of step t of a case (Channel::step, src/cases/channel.rs:136-147): the inflow ramp,
// the body kick (host-tabulated), the outflow target.
__device__ inline StepCtl case_controls(const BatchCase& cs, int64_t t, const double* __restrict__ kick_tab) {
  StepCtl ctl;
  const double r = ramp_dev((uint64_t)t, cs.ramp_steps);
  ctl.vscale[0] = 1.0;
  ctl.vscale[1] = r;
  ctl.vscale[2] = (t >= cs.kick_start && t < cs.kick_start + cs.kick_dur) ? kick_tab[cs.kick_off + t - cs.kick_start]
                                                                         : 0.0;
  ctl.vscale[3] = 1.0;
  ctl.target = cs.u_ref * r;
  ctl.has_target = cs.absorbing && cs.t.np > 0;
  return ctl;
}

template <class T>
__device__ inline BandView<T> band_view(const BatchCase& cs, T* pool, T* side) {
  return BandView<T>{pool + cs.pop_off, cs.g.plane, cs.g.pitch, cs.band_w > 0 && side ? side + cs.side_off : nullptr,
                     cs.band_w, cs.g.nx - 1 - cs.band_w, cs.band_sny};
}

// Boundary tiles (256 boundary nodes of one case each) of one step: the tiles of the cases still
// running, [first, act[t]), where first = 0 in single steps; in step `jpos` of a group of K steps
// that starts at t - jpos, first = first_at[t - jpos] skips the tiles of the cases whose boundary
// nodes the region kernel advances.  The first tile of a case also reduces the previous step's
// force terms into its history.
template <class S>
__global__ __launch_bounds__(BT) void k_batch_boundary(const typename S::T* __restrict__ A,
                                                       typename S::T* __restrict__ B,
                                                       const BatchCase* __restrict__ cases,
                                                       const int32_t* __restrict__ tcase,
                                                       const int32_t* __restrict__ tb0,
                                                       const int32_t* __restrict__ act, const int64_t* tbase,
                                                       int t_local, const double* __restrict__ kick_tab,
                                                       double* __restrict__ hist, int32_t* __restrict__ status,
                                                       int64_t* __restrict__ done, double sqrt_cs2,
                                                       const int32_t* __restrict__ first_at, int jpos,
                                                       const int32_t* __restrict__ band_end_at,
                                                       const typename S::T* __restrict__ side_rd) {
  __shared__ double red[BT];
  __shared__ double fsum[2];
  __shared__ int skip;
  const int64_t t = *tbase + t_local;
  const int nact = act[t];
  const int first = first_at ? first_at[t - jpos] : 0;
  // band_end_at (steps of a group with side arrays): the tiles below band_end_at[t - jpos] belong
  // to the cases in passes, whose band nodes the band kernel advances
  const int band_end = band_end_at ? band_end_at[t - jpos] : 0;
  for (int i = first + (int)blockIdx.x; i < nact; i += gridDim.x) {
    const int ci = tcase[i];
    if (threadIdx.x == 0) skip = status[ci];
    __syncthreads();
    const int sk = skip;
    __syncthreads();
    if (sk) continue;
    const BatchCase& cs = cases[ci];
    const int64_t ns2 = (int64_t)max(cs.t.nslots, 1) * 2;
    const int64_t np1 = max(cs.t.np, 1);
    // force terms of the previous step -> coefficient history (and the divergence check)
    if (i == cs.bnd_tile0 && t >= 1 && t - 1 < cs.steps) {
      double cd = 0.0, cl = 0.0;
      if (cs.t.ngroups > 0) {
        reduce_forces(cs.t.fcontrib + ((t + 1) & 1) * ns2, cs.t.gfirst, 1, fsum, red);
        cd = cs.coef_k * fsum[0];
        cl = cs.coef_k * fsum[1];
      }
      if (threadIdx.x == 0) {
        hist[cs.hist_off + 2 * (t - 1)] = cd;
        hist[cs.hist_off + 2 * (t - 1) + 1] = cl;
        done[ci] = t;
        if (!isfinite(cd)) status[ci] = 1;
      }
    }
    if (t < cs.steps) {
      const StepCtl ctl = case_controls(cs, t, kick_tab);
      const int b0 = tb0[i];
      double drho = 0.0;
      if (ctl.has_target && b0 < cs.t.npb)
        drho = outlet_offset<S>(cs.t.pvel + (t & 1) * np1, cs.t.np, ctl.target, sqrt_cs2, red);
      if (b0 == 0 && threadIdx.x == 0) *cs.t.drho = drho;
      const int b = b0 + (int)threadIdx.x;
      const bool band = i < band_end && cs.band_w > 0;
      if (b < cs.t.nb) {
        using T = typename S::T;
        const int bx = cs.t.bx[b];
        const bool skip_node = band && (bx <= cs.band_w || bx >= cs.g.nx - 1 - cs.band_w);
        if (!skip_node) {
          const BandView<const T> rv = band_view<const T>(cs, A, band ? side_rd : nullptr);
          const MainView<T> wv{B + cs.pop_off, cs.g.plane, cs.g.pitch};
          double* fc = cs.t.fcontrib + (t & 1) * ns2;
          double* pvn = cs.t.pvel + ((t + 1) & 1) * np1;
          if (cs.ls.meta) {
            if (cs.c.model == MODEL_BGK)
              boundary_node_v<S, MODEL_BGK, true>(rv, wv, cs.g, cs.c, cs.t, ctl, drho, sqrt_cs2, b, fc, pvn, &cs.ls);
            else
              boundary_node_v<S, MODEL_TRT, true>(rv, wv, cs.g, cs.c, cs.t, ctl, drho, sqrt_cs2, b, fc, pvn, &cs.ls);
          } else if (cs.c.model == MODEL_BGK) {
            boundary_node_v<S, MODEL_BGK>(rv, wv, cs.g, cs.c, cs.t, ctl, drho, sqrt_cs2, b, fc, pvn);
          } else {
            boundary_node_v<S, MODEL_TRT>(rv, wv, cs.g, cs.c, cs.t, ctl, drho, sqrt_cs2, b, fc, pvn);
          }
        }
      }
    }
    __syncthreads();  // `red` is reused by the next tile
  }
}

// Bulk tiles: TW x TH interior nodes of one case per pass of a workgroup (TW * TH = BT).  The
// tiles run are [first, act[t]), where first = 0 in single steps; in step `jpos` of a group of
// K steps that starts at t - jpos, first = first_at[t - jpos] skips the tiles of the cases that
// run the whole group in K-step passes.
template <class S>
__global__ __launch_bounds__(BT) void k_batch_bulk(const typename S::T* __restrict__ A, typename S::T* __restrict__ B,
                                                   const uint8_t* __restrict__ flags,
                                                   const BatchCase* __restrict__ cases,
                                                   const int32_t* __restrict__ tcase, const int32_t* __restrict__ ty0,
                                                   const int32_t* __restrict__ tx0, const int32_t* __restrict__ act,
                                                   const int64_t* tbase, int t_local, int tw_log2,
                                                   const int32_t* __restrict__ status,
                                                   const int32_t* __restrict__ first_at, int jpos) {
  const int64_t t = *tbase + t_local;
  const int nact = act[t];
  const int first = first_at ? first_at[t - jpos] : 0;
  const int tx = (int)threadIdx.x & ((1 << tw_log2) - 1);
  const int ty = (int)threadIdx.x >> tw_log2;
  for (int i = first + (int)blockIdx.x; i < nact; i += gridDim.x) {
    const int ci = tcase[i];
    if (status[ci]) continue;
    const BatchCase& cs = cases[ci];
    const int x = tx0[i] + tx, y = ty0[i] + ty;
    if (x >= cs.g.nx || y >= cs.g.ny) continue;
    const int64_t idx = (int64_t)y * cs.g.pitch + x;
    if (flags[cs.flag_off + idx] != 0) continue;
    const typename S::T* a = A + cs.pop_off;
    typename S::T* b = B + cs.pop_off;
    typename S::C f[Q];
    pull<S>(a, cs.g, x, y, f);
    if (cs.c.model == MODEL_BGK)
      collide<S, MODEL_BGK>(f, cs.c);
    else
      collide<S, MODEL_TRT>(f, cs.c);
#pragma unroll
    for (int q = 0; q < Q; ++q) S::st_stream(b + q * cs.g.plane + idx, f[q]);
  }
}

// K steps (t .. t+K) of the covered interior of the cases that run the whole group: one task
// per wave, a strip of one case (task_case), on the case's arrays at its offset in each pool.
// The cases are in descending step order, so those tasks are a prefix of ntask_at[t] tasks.
// The wave-uniform branch on the case's model inlines the strip once per collide; it needs
// more registers (3 waves per SIMD for FP32 at K = 5) than a branch inside each collide (4),
// and still ran the 512-case sweep 5 % faster.
template <class S, int W, int V, int D, int K>
__global__ __launch_bounds__(W * TB_LANES) void k_batch_pass(const typename S::T* __restrict__ A,
                                                            PassOut<typename S::T> po, const int4* __restrict__ tasks,
                                                            const int32_t* __restrict__ task_case,
                                                            const int32_t* __restrict__ ntask_at, const int64_t* tbase,
                                                            int t_local, const BatchCase* __restrict__ cases,
                                                            const int32_t* __restrict__ status) {
  using T = typename S::T;
  const int64_t t = *tbase + t_local;
  const int w = __builtin_amdgcn_readfirstlane((int)blockIdx.x * W + (int)(threadIdx.x / TB_LANES));
  if (w >= ntask_at[t]) return;
  const int4 task = tasks[w];
  const int ci = __builtin_amdgcn_readfirstlane(task_case[w]);
  // A workgroup in lockstep holds the tasks of one case only (the prefix ends between cases, so
  // all its waves got here): one read of the case's status serves them all, so they all run and
  // meet the same barriers, or all return.
  int skip;
  if ((task.w >> 12) & 1) {
    __shared__ int sk;
    if (threadIdx.x == 0) sk = status[ci];
    __syncthreads();
    skip = sk;
  } else {
    skip = status[ci];
  }
  if (skip) return;
  const BatchCase& cs = cases[ci];
  const int64_t off = cs.pop_off;
  PassOut<T> pc;
#pragma unroll
  for (int L = 0; L < TB_MAX_K - 1; ++L) pc.ring[L] = L < K - 1 ? po.ring[L] + off : nullptr;
  pc.out = po.out + off;
  if (cs.c.model == MODEL_BGK)
    multi_step_strip<S, MODEL_BGK, V, D, K>(task, A + off, pc, cs.g, cs.c);
  else
    multi_step_strip<S, MODEL_TRT, V, D, K>(task, A + off, pc, cs.g, cs.c);
}

// One single step of the interior nodes that no task covers, for the cases in K-step passes:
// listed tile i holds count[i] <= LISTED_NPT * BT nodes of case lcase[i] (packed y << 16 | x from
// nodes + first[i]); the tiles of the cases that run the group starting at t - jpos are a prefix
// of nlist_at[t - jpos] tiles.  With side arrays the lists hold no band node, and the band
// neighbours of time t are read from side_rd (null at a group's first step: the pool).
template <class S>
__global__ __launch_bounds__(BT) void k_batch_listed(const typename S::T* __restrict__ A, typename S::T* __restrict__ B,
                                                     const BatchCase* __restrict__ cases,
                                                     const int32_t* __restrict__ lcase, const int32_t* __restrict__ lfirst,
                                                     const int32_t* __restrict__ lcount,
                                                     const uint32_t* __restrict__ nodes,
                                                     const int32_t* __restrict__ nlist_at, const int64_t* tbase,
                                                     int t_local, int jpos, const int32_t* __restrict__ status,
                                                     const typename S::T* __restrict__ side_rd) {
  using T = typename S::T;
  using C = typename S::C;
  const int64_t t = *tbase + t_local;
  const int n = nlist_at[t - jpos];
  for (int i = blockIdx.x; i < n; i += gridDim.x) {
    const int ci = lcase[i];
    if (status[ci]) continue;
    const BatchCase& cs = cases[ci];
    const int cnt = lcount[i];
    const BandView<const T> rv = band_view<const T>(cs, A, side_rd);
    const MainView<T> wv{B + cs.pop_off, cs.g.plane, cs.g.pitch};
    // LISTED_NPT nodes per thread (k, k + BT, ...): every node's pulls are issued before the
    // first collide, so a wave has that many nodes' loads in flight
    int x[LISTED_NPT], y[LISTED_NPT];
    bool on[LISTED_NPT];
    C f[LISTED_NPT][Q];
#pragma unroll
    for (int m = 0; m < LISTED_NPT; ++m) {
      const int k = (int)threadIdx.x + m * BT;
      on[m] = k < cnt;
      const uint32_t v = on[m] ? nodes[lfirst[i] + k] : 0u;
      x[m] = (int)(v & 0xffffu);
      y[m] = (int)(v >> 16);
      if (on[m]) pull_v<S>(rv, cs.g, x[m], y[m], f[m]);
    }
#pragma unroll
    for (int m = 0; m < LISTED_NPT; ++m) {
      if (!on[m]) continue;
      if (cs.c.model == MODEL_BGK)
        collide<S, MODEL_BGK>(f[m], cs.c);
      else
        collide<S, MODEL_TRT>(f[m], cs.c);
#pragma unroll
      for (int q = 0; q < Q; ++q) S::st(wv.at(q, x[m], y[m]), f[m][q]);
    }
  }
}

// One single step of the band nodes of the cases in K-step passes (see BandView): band tile i =
// (case, band column, first row) covers BT rows of one band column, a node per thread, so each
// load and store of a side array runs along a line.  bidx maps a band position to -2 (not a
// fluid node), -1 (interior) or the boundary node's index.  Step jpos of the group reads the
// band's time t from side_rd (the pool at the group's first step) and writes time t + 1 to
// side_wr (the pool at its last step).  A tile of the outlet column computes the outlet
// controller's offset as the boundary kernel's tiles do.
template <class S>
__global__ __launch_bounds__(BT) void k_batch_band(const typename S::T* __restrict__ A, typename S::T* __restrict__ B,
                                                   const typename S::T* __restrict__ side_rd,
                                                   typename S::T* __restrict__ side_wr,
                                                   const BatchCase* __restrict__ cases, const int4* __restrict__ tiles,
                                                   const int32_t* __restrict__ nband_at,
                                                   const int32_t* __restrict__ bidx, const int64_t* tbase,
                                                   int t_local, int jpos, const double* __restrict__ kick_tab,
                                                   const int32_t* __restrict__ status, double sqrt_cs2) {
  using T = typename S::T;
  __shared__ double red[BT];
  __shared__ int skip;
  const int64_t t = *tbase + t_local;
  const int n = nband_at[t - jpos];
  for (int i = blockIdx.x; i < n; i += gridDim.x) {
    const int4 tl = tiles[i];  // case, band column, first row
    const int ci = tl.x;
    if (threadIdx.x == 0) skip = status[ci];
    __syncthreads();
    const int sk = skip;
    __syncthreads();
    if (sk) continue;
    const BatchCase& cs = cases[ci];
    const int bw = cs.band_w, xr = cs.g.nx - 1 - bw;
    const int c = tl.y;
    const int x = c < bw ? 1 + c : xr + (c - bw);
    const int y = tl.z + (int)threadIdx.x;
    const StepCtl ctl = case_controls(cs, t, kick_tab);
    const int64_t np1 = max(cs.t.np, 1);
    double drho = 0.0;
    if (ctl.has_target && x == cs.g.nx - 2)  // the outlet column: its nodes use the controller
      drho = outlet_offset<S>(cs.t.pvel + (t & 1) * np1, cs.t.np, ctl.target, sqrt_cs2, red);
    const int32_t b = y < cs.band_sny ? bidx[cs.bidx_off + (int64_t)c * cs.band_sny + y] : -2;
    if (b != -2) {
      const BandView<const T> rv = band_view<const T>(cs, A, side_rd);
      const BandView<T> wv = band_view<T>(cs, B, side_wr);
      if (b >= 0) {
        const int64_t ns2 = (int64_t)max(cs.t.nslots, 1) * 2;
        double* fc = cs.t.fcontrib + (t & 1) * ns2;
        double* pvn = cs.t.pvel + ((t + 1) & 1) * np1;
        if (cs.ls.meta) {
          if (cs.c.model == MODEL_BGK)
            boundary_node_v<S, MODEL_BGK, true>(rv, wv, cs.g, cs.c, cs.t, ctl, drho, sqrt_cs2, b, fc, pvn, &cs.ls);
          else
            boundary_node_v<S, MODEL_TRT, true>(rv, wv, cs.g, cs.c, cs.t, ctl, drho, sqrt_cs2, b, fc, pvn, &cs.ls);
        } else if (cs.c.model == MODEL_BGK) {
          boundary_node_v<S, MODEL_BGK>(rv, wv, cs.g, cs.c, cs.t, ctl, drho, sqrt_cs2, b, fc, pvn);
        } else {
          boundary_node_v<S, MODEL_TRT>(rv, wv, cs.g, cs.c, cs.t, ctl, drho, sqrt_cs2, b, fc, pvn);
        }
      } else {
        if (cs.c.model == MODEL_BGK)
          interior_node_v<S, MODEL_BGK>(x, y, rv, wv, cs.g, cs.c);
        else
          interior_node_v<S, MODEL_TRT>(x, y, rv, wv, cs.g, cs.c);
      }
    }
    __syncthreads();  // `red` is reused by the next tile
  }
}

// The slot-major copy of every case's links (built once, see LinkSlots): workgroup row y takes
// case case0 + y, a boundary node per thread.
__global__ __launch_bounds__(BT) void k_link_slots(const BatchCase* __restrict__ cases, int case0, uint32_t* meta,
                                                   int32_t* fslot, double* q, double* cu6, double* rho, double* uwx,
                                                   double* uwy) {
  const BatchCase& cs = cases[case0 + (int)blockIdx.y];
  if (!cs.ls.meta) return;
  const DevTables& t = cs.t;
  const int64_t stride = cs.ls.stride;
  const int64_t base = cs.ls.meta - meta;  // the case's first boundary node in the slot arrays
  for (int b = (int)(blockIdx.x * BT + threadIdx.x); b < t.nb; b += (int)(gridDim.x * BT)) {
    const int l0 = t.bfirst[b], n = min(t.bfirst[b + 1] - l0, LINK_SLOTS);
    for (int k = 0; k < max(n, 1); ++k) {
      const int64_t e = base + (int64_t)k * stride + b;
      if (k >= n) {  // a node without links: slot 0 says so
        meta[e] = pack_link_meta(0, 0, 0, 0, 0, 0, 0);
        continue;
      }
      const int l = l0 + k;
      meta[e] = pack_link_meta(t.dir[l], t.kind[l], t.vgroup[l], t.lflags[l], t.lnx[l], t.lny[l], k == 0 ? n : 0);
      fslot[e] = t.fslot[l];
      q[e] = t.q[l];
      cu6[e] = t.cu6[l];
      rho[e] = t.rho[l];
      uwx[e] = t.uwx[l];
      uwy[e] = t.uwy[l];
    }
  }
}

// The band map of every case (built once): -2 for a non-fluid band position, -1 for an interior
// node, then (second kernel) the index of each boundary node that lies in a band.
__global__ __launch_bounds__(BT) void k_band_map_init(int32_t* __restrict__ bidx, const uint8_t* __restrict__ flags,
                                                      const BatchCase* __restrict__ cases, int case0) {
  const BatchCase& cs = cases[case0 + (int)blockIdx.y];
  if (cs.band_w <= 0) return;
  const int64_t n = (int64_t)2 * cs.band_w * cs.band_sny;
  for (int64_t e = (int64_t)blockIdx.x * BT + threadIdx.x; e < n; e += (int64_t)gridDim.x * BT) {
    const int c = (int)(e / cs.band_sny), y = (int)(e - (int64_t)c * cs.band_sny);
    const int x = c < cs.band_w ? 1 + c : cs.g.nx - 1 - cs.band_w + (c - cs.band_w);
    bidx[cs.bidx_off + e] = y < cs.g.ny && flags[cs.flag_off + (int64_t)y * cs.g.pitch + x] != 2 ? -1 : -2;
  }
}
__global__ __launch_bounds__(BT) void k_band_map_bnd(int32_t* __restrict__ bidx, const BatchCase* __restrict__ cases,
                                                     int case0) {
  const BatchCase& cs = cases[case0 + (int)blockIdx.y];
  if (cs.band_w <= 0) return;
  const int xr = cs.g.nx - 1 - cs.band_w;
  for (int b = (int)(blockIdx.x * BT + threadIdx.x); b < cs.t.nb; b += (int)(gridDim.x * BT)) {
    const int x = cs.t.bx[b], y = cs.t.by[b];
    const bool lo = x >= 1 && x <= cs.band_w, hi = x >= xr && x < xr + cs.band_w;
    if (lo || hi) bidx[cs.bidx_off + (int64_t)(lo ? x - 1 : cs.band_w + x - xr) * cs.band_sny + y] = b;
  }
}

// The bands' level K of a group, which its last step wrote into the level-0 side array, copied
// into the pool (the next group's passes read it there as halo, and single steps outside groups
// read every node there).  Copy tile i = (case, first row) covers COPY_ROWS rows of all 2 band_w
// band columns: the workgroup loads the block of every plane along the side array's columns (the
// rows of a column on adjacent threads), turns it over in LDS, and stores it along the pool's rows
// (the columns of a row on adjacent threads, two runs of band_w values per row).  Positions that
// are not fluid keep their pool values.
constexpr int COPY_ROWS = 16;
template <class T>
__global__ __launch_bounds__(BT) void k_band_to_pool(T* __restrict__ B, const T* __restrict__ side0,
                                                     const BatchCase* __restrict__ cases, const int2* __restrict__ tiles,
                                                     const int32_t* __restrict__ ncopy_at,
                                                     const int32_t* __restrict__ bidx, const int64_t* tbase,
                                                     int t_local, int jpos, const int32_t* __restrict__ status) {
  constexpr int NE = COPY_ROWS * 2 * TB_MAX_K;  // block elements per plane, at most
  __shared__ T blk[Q * NE];
  __shared__ uint8_t fluid[NE];
  __shared__ int skip;
  const int64_t t = *tbase + t_local;
  const int n = ncopy_at[t - jpos];
  for (int i = blockIdx.x; i < n; i += gridDim.x) {
    const int2 tl = tiles[i];  // case, first row
    const int ci = tl.x;
    // one read of the case's status for the whole workgroup (it meets barriers below)
    if (threadIdx.x == 0) skip = status[ci];
    __syncthreads();
    const int sk = skip;
    __syncthreads();
    if (sk) continue;
    const BatchCase& cs = cases[ci];
    const int bw = cs.band_w, bw2 = 2 * bw, ne = COPY_ROWS * bw2;
    // load: element e = column c, row r (rows fastest); the block is stored row-major
    for (int e = (int)threadIdx.x; e < ne; e += BT) {
      const int c = e / COPY_ROWS, r = e - c * COPY_ROWS, y = tl.y + r;
      const int64_t col = (int64_t)c * cs.band_sny + y;
      const bool ok = y < cs.g.ny && bidx[cs.bidx_off + col] != -2;
      const int o = r * bw2 + c;
      fluid[o] = ok;
      if (ok) {
        const T* src = side0 + cs.side_off + col;
#pragma unroll
        for (int k = 0; k < Q; ++k) blk[k * NE + o] = src[(int64_t)k * bw2 * cs.band_sny];
      }
    }
    __syncthreads();
    // store: element e = row r, column c (columns fastest)
    for (int e = (int)threadIdx.x; e < ne; e += BT) {
      if (!fluid[e]) continue;
      const int r = e / bw2, c = e - r * bw2, y = tl.y + r;
      const int x = c < bw ? 1 + c : cs.g.nx - 1 - bw + (c - bw);
      T* dst = B + cs.pop_off + (int64_t)y * cs.g.pitch + x;
#pragma unroll
      for (int k = 0; k < Q; ++k) dst[(int64_t)k * cs.g.plane] = blk[k * NE + e];
    }
    __syncthreads();  // the block is reused by the next tile
  }
}

// The K + 1 population pools of a group of K steps, in time order (pool j holds time t + j).
template <class T>
struct PoolSet {
  T* p[TB_MAX_K + 1];
};

// Steps j0 .. j1 - 1 of a group of K steps from t, for the boundary nodes and the listed interior
// nodes of the cases that run the whole group in K-step passes: one workgroup per case (the
// cases with steps >= t + K, rcase[0 .. nreg_at[t])), a barrier between steps, so the steps of a
// case need no launch each.  Step j reads pool j and writes pool j + 1; from j = 1 on it reads
// the ring values of level j that the pass writes, so such a launch follows the pass.  Per node
// the arithmetic is the boundary and listed kernels', and the per-step work of the boundary
// kernel's first tile (the previous step's force reduction, the history, the divergence check,
// the controls and the outlet controller) happens once per step with the same fixed-order
// trees, so a case gives the same bits as in single steps.
template <class S, int RT>
__global__ __launch_bounds__(RT) void k_batch_region(PoolSet<typename S::T> ps, const BatchCase* __restrict__ cases,
                                                     const int32_t* __restrict__ rcase,
                                                     const int32_t* __restrict__ nreg_at,
                                                     const int32_t* __restrict__ clfirst,
                                                     const int32_t* __restrict__ clcount,
                                                     const uint32_t* __restrict__ lnodes, const int64_t* tbase,
                                                     int t_local, int j0, int j1, const double* __restrict__ kick_tab,
                                                     double* __restrict__ hist, int32_t* __restrict__ status,
                                                     int64_t* __restrict__ done, double sqrt_cs2) {
  __shared__ double red[STEP_T];
  __shared__ double fsum[2];
  __shared__ int skip;
  const int64_t t = *tbase + t_local;  // the group's first step
  const int n = nreg_at[t];
  for (int r = blockIdx.x; r < n; r += gridDim.x) {
    const int ci = rcase[r];
    const BatchCase& cs = cases[ci];
    const int64_t ns2 = (int64_t)max(cs.t.nslots, 1) * 2;
    const int64_t np1 = max(cs.t.np, 1);
    const uint32_t* ln = lnodes + clfirst[ci];
    const int nln = clcount[ci];
    for (int j = j0; j < j1; ++j) {
      const int64_t tj = t + j;  // < cs.steps: the case runs the whole group
      if (threadIdx.x == 0) skip = status[ci];
      __syncthreads();
      const int sk = skip;
      __syncthreads();
      if (sk) break;
      // force terms of step tj - 1 -> coefficient history (and the divergence check)
      if (tj >= 1) {
        double cd = 0.0, cl = 0.0;
        if (cs.t.ngroups > 0) {
          reduce_forces<RT>(cs.t.fcontrib + ((tj + 1) & 1) * ns2, cs.t.gfirst, 1, fsum, red);
          cd = cs.coef_k * fsum[0];
          cl = cs.coef_k * fsum[1];
        }
        if (threadIdx.x == 0) {
          hist[cs.hist_off + 2 * (tj - 1)] = cd;
          hist[cs.hist_off + 2 * (tj - 1) + 1] = cl;
          done[ci] = tj;
          if (!isfinite(cd)) status[ci] = 1;
        }
      }
      StepCtl ctl;
      const double rmp = ramp_dev((uint64_t)tj, cs.ramp_steps);
      ctl.vscale[0] = 1.0;
      ctl.vscale[1] = rmp;
      ctl.vscale[2] = (tj >= cs.kick_start && tj < cs.kick_start + cs.kick_dur) ? kick_tab[cs.kick_off + tj - cs.kick_start]
                                                                                : 0.0;
      ctl.vscale[3] = 1.0;
      ctl.target = cs.u_ref * rmp;
      ctl.has_target = cs.absorbing && cs.t.np > 0;
      double drho = 0.0;
      if (ctl.has_target) drho = outlet_offset<S, RT>(cs.t.pvel + (tj & 1) * np1, cs.t.np, ctl.target, sqrt_cs2, red);
      if (threadIdx.x == 0) *cs.t.drho = drho;
      const typename S::T* a = ps.p[j] + cs.pop_off;
      typename S::T* bb = ps.p[j + 1] + cs.pop_off;
      double* fc = cs.t.fcontrib + (tj & 1) * ns2;
      double* pvn = cs.t.pvel + ((tj + 1) & 1) * np1;
      // (only the first npb boundary nodes have links that use drho: the outlet-controller
      // nodes come first)
      if (cs.c.model == MODEL_BGK) {
        for (int b = (int)threadIdx.x; b < cs.t.nb; b += RT)
          boundary_node<S, MODEL_BGK>(a, bb, cs.g, cs.c, cs.t, ctl, drho, sqrt_cs2, b, fc, pvn);
        for (int k = (int)threadIdx.x; k < nln; k += RT) listed_node<S, MODEL_BGK>(k, nln, ln, a, bb, cs.g, cs.c);
      } else {
        for (int b = (int)threadIdx.x; b < cs.t.nb; b += RT)
          boundary_node<S, MODEL_TRT>(a, bb, cs.g, cs.c, cs.t, ctl, drho, sqrt_cs2, b, fc, pvn);
        for (int k = (int)threadIdx.x; k < nln; k += RT) listed_node<S, MODEL_TRT>(k, nln, ln, a, bb, cs.g, cs.c);
      }
      __syncthreads();  // step tj written before step tj + 1 reads it
    }
  }
}

// Initial state of every case: rest, or the inflow profile when the case has no ramp
// (Channel::new, src/cases/channel.rs:126-128); non-fluid cells at rest (src/sim.rs:362).
template <class S>
__global__ __launch_bounds__(BT) void k_batch_init(typename S::T* __restrict__ A, typename S::T* __restrict__ B,
                                                   const uint8_t* __restrict__ flags,
                                                   const BatchCase* __restrict__ cases, int case0) {
  const BatchCase& cs = cases[case0 + (int)blockIdx.y];  // one row of workgroups per case
  const int64_t n = (int64_t)cs.g.pitch * cs.g.ny;
  for (int64_t e = (int64_t)blockIdx.x * BT + threadIdx.x; e < n; e += (int64_t)gridDim.x * BT) {
    const int y = (int)(e / cs.g.pitch), x = (int)(e - (int64_t)y * cs.g.pitch);
    double ux = 0.0;
    if (x < cs.g.nx && flags[cs.flag_off + e] != 2 && cs.ramp_steps == 0) {
      if (cs.inflow == VL_INFLOW_PARABOLIC) {
        const double h = cs.inflow_h, u_mean = cs.inflow_u;
        const double s = fmin(fmax((double)y - 0.5, 0.0), h);
        ux = 6.0 * u_mean * s * (h - s) / (h * h);
      } else {
        ux = cs.inflow_u;
      }
    }
    double feq[Q];
    Exact::equilibrium(1.0, ux, 0.0, feq);
#pragma unroll
    for (int q = 0; q < Q; ++q) {
      typename S::C v;
      if constexpr (S::exact)
        v = feq[q];
      else
        v = (float)(feq[q] - wgt(q));
      S::st(A + cs.pop_off + q * cs.g.plane + e, v);
      S::st(B + cs.pop_off + q * cs.g.plane + e, v);
    }
  }
}

// Outward velocities of every case's outlet-controller nodes from its populations in pool A (start
// of a run): k_outlet_prime of the single domain, one row of workgroups per case.
template <class S>
__global__ __launch_bounds__(BT) void k_batch_outlet_prime(const typename S::T* __restrict__ A,
                                                           const BatchCase* __restrict__ cases, int case0) {
  const BatchCase& cs = cases[case0 + (int)blockIdx.y];
  const DevTables& t = cs.t;
  for (int k = (int)(blockIdx.x * BT + threadIdx.x); k < t.np; k += (int)(gridDim.x * BT)) {
    const int64_t idx = (int64_t)t.py[k] * cs.g.pitch + t.px[k];
    typename S::C f[Q];
#pragma unroll
    for (int i = 0; i < Q; ++i) f[i] = S::ld(A + cs.pop_off + i * cs.g.plane + idx);
    t.pvel[k] = normal_velocity_of<S>(f, cs.c, t.pnx[k], t.pny[k]);
  }
}

__global__ void k_batch_advance(int64_t* tbase, int64_t by) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *tbase += by;
}

}  // namespace vl

using namespace vl;

// ==========================================================================================
// Host side
// ==========================================================================================
struct vl_batch {
  int precision = VL_F32;
  size_t elem = 4;
  struct Pending {
    vl_channel_spec spec;
    vl_shape* body;
    int64_t steps;
    double d;
  };
  std::vector<Pending> pending;
  bool built = false;
  // per case (host)
  std::vector<BatchCase> hcases;
  std::vector<Tables> tabs;  // with device set-up: only nx, ny, fluid_nodes and flag are filled
  int64_t total_nodes = 0, total_fluid = 0, max_steps = 0;
  bool setup_on_device = false;  // cell kinds, flags and boundary tables built by the set-up kernels
  double setup_ms[4] = {};      // set-up time: cases and tables, pass plans, packing and uploads, autotuning
  // device
  hipStream_t stream = nullptr, side = nullptr;
  hipEvent_t ev_fork = nullptr, ev_join = nullptr, ev_mid = nullptr;
  // population pools: 2, or K + 1 with K-step passes (pool j + 1 follows pool j in time)
  void* pool[TB_MAX_K + 1] = {};
  int nbuf = 2;
  // K-step passes (pass_k > 1): the covered interior of a case that runs a whole group of K
  // steps advances in one pass; its other interior nodes are listed single steps.  Tasks and
  // listed tiles are in the cases' descending step order.
  int pass_k = 1, pass_rows = 0, pass_waves = 8;
  int nlisted_tiles = 0;
  int64_t pass_nodes = 0, listed_nodes = 0;
  // Pass tasks, each case's together (BGK and TRT cases alike: the pass kernel branches on
  // the case's model).  Per group start t, n_at[t] = the tasks of the cases with steps >= t + K
  // (a prefix).
  struct TaskList {
    int4* task = nullptr;
    int32_t* tcase = nullptr;
    int32_t* n_at = nullptr;
    int n = 0;
    std::vector<int32_t> h_n_at;
  } tasks, tasks_far, tasks_near;
  // with the region kernel: the region kernel of a group runs beside the next group's far pass
  // (VL_BATCH_OVERLAP=0: after the group's whole pass, before the next group)
  bool overlap = true;
  bool ov_open = false;  // the side stream runs ahead of the main stream (joined by join_side)
  int32_t *d_kfirst_at = nullptr;
  int32_t *d_lcase = nullptr, *d_lfirst = nullptr, *d_lcount = nullptr, *d_ln_at = nullptr;
  uint32_t* d_lnodes = nullptr;
  // per group start t: listed tiles of the cases with steps >= t + K, and the bulk tiles of
  // those cases (the bulk kernel skips them in a group)
  std::vector<int32_t> h_ln_at, h_kfirst_at;
  // Region kernel (VL_BATCH_REGION=1, with K-step passes; by default the boundary and listed
  // kernels run once per step): one workgroup of region_threads per case in passes advances its
  // boundary and listed nodes through the group's steps.  rcase = the cases in descending step
  // order; per group start t, rn_at[t] = the cases with steps >= t + K, bfirst_at[t] = their
  // boundary tiles (a prefix, which the boundary kernel then skips).  clfirst / clcount: each
  // case's listed nodes in d_lnodes.  region0_side: step 0 of the region kernel beside the pass.
  bool region = false, region0_side = true;
  // Side arrays of the bands (see BandView; the default with K-step passes, VL_BATCH_BAND=0
  // keeps every node in the pools): bside[L - 1] holds the bands' level L = 1 .. K-1 of a group,
  // bidx the band map, band tiles (case, band column, first row) in the cases' descending step
  // order, nband_at[t] = the band tiles of the cases in passes for the group starting at t.
  // bside0 holds the bands' level K of the last group (the next group's level 0), written
  // beside the pool at a group's last step; side0_ok says that it matches the pools (no single
  // step ran since), so a group's first step reads its bands there.  Copy tiles (case, first
  // row; ncopy_at[t] per group start like nband_at) write level K into the pool row by row.
  bool band = false;
  void* bside[TB_MAX_K - 1] = {};
  void* bside0 = nullptr;
  bool side0_ok = false;
  int32_t* d_bidx = nullptr;
  int4* d_btile = nullptr;
  int32_t* d_nband_at = nullptr;
  std::vector<int32_t> h_nband_at;
  int nband_tiles = 0;
  int2* d_ctile = nullptr;
  int32_t* d_ncopy_at = nullptr;
  std::vector<int32_t> h_ncopy_at;
  int ncopy_tiles = 0;
  // Slot-major copy of the links (LinkSlots) for the boundary and band kernels; VL_BATCH_LINK_SLOTS=0
  // keeps them on the link tables' columns.  Arrays: meta, fslot, q, cu6, rho, uwx, uwy.
  bool link_slots = true;
  void* lslot[7] = {};
  hipStream_t third = nullptr;  // the band kernel's stream in the steps after a group's first
  hipEvent_t ev_join3 = nullptr;
  // Single steps beside the next group's far pass (VL_BATCH_STEP_OVERLAP=1, with side arrays and
  // a third stream): steps 1 .. K-1 of a group run on the side stream, forked onto the third
  // (band) and fourth (listed and bulk) streams, while the main stream runs the next group's far
  // tasks.  Bit-identical, but on the 512-case sweep 5 % slower than the default (the single
  // steps after the group's whole pass): the far pass beside them slows by about as much time
  // as they take.
  bool step_overlap = false;
  hipStream_t fourth = nullptr;
  hipEvent_t ev_sfork = nullptr, ev_join4 = nullptr;
  int region_threads = 512;
  int32_t *d_rcase = nullptr, *d_rn_at = nullptr, *d_bfirst_at = nullptr, *d_clfirst = nullptr,
          *d_clcount = nullptr;
  std::vector<int32_t> h_rn_at, h_bfirst_at, h_bact;
  uint8_t* dflags = nullptr;
  BatchCase* dcases = nullptr;
  std::vector<void*> allocs;  // table pools etc.
  int32_t *d_bcase = nullptr, *d_bb0 = nullptr, *d_bact = nullptr;
  int32_t *d_kcase = nullptr, *d_ky0 = nullptr, *d_kx0 = nullptr, *d_kact = nullptr;
  int32_t* d_status = nullptr;
  int64_t* d_done = nullptr;
  int64_t* d_tbase = nullptr;  // [0] window counter, [1] a zero
  double* d_hist = nullptr;
  double* d_kick = nullptr;
  int nbnd_tiles = 0, nbulk_tiles = 0;
  int tw_log2 = 6, th = 4, grid_bulk = 0, grid_bnd = 0;
  // bulk grid policy: 0 = one workgroup per active tile (the grid follows the active-tile count
  // on a geometric ladder), m > 0 = persistent grid of m workgroups per CU
  int bulk_mult = 0;
  bool serial = false;  // boundary kernel before the bulk on one stream (else the two run side by side)
  std::vector<int32_t> h_kact;  // active bulk tiles per step (host copy of d_kact)
  std::vector<int32_t> ladder;  // grid sizes used when bulk_mult == 0
  int cur = 0;
  int64_t t = 0;
  // (first pool, bulk grid, pass grid, listed or region grid, boundary grid, near pass grid,
  // band grid, copy grid, side0_ok) -> window
  std::map<std::tuple<int, int, int, int, int, int, int, int, int>, hipGraphExec_t> graphs;
  bool use_graphs = true;
  double autotune_ms[3][4] = {};
  double last_run_seconds = 0.0;
  int64_t device_bytes = 0;
  // kernel timing (vl_batch_set_timing): events around each kernel of each step, read back
  // every TIMING_W steps
  struct StepEvents {
    hipEvent_t bs, be, ks, ke, ej;
  };
  bool timing = false;
  std::vector<StepEvents> tev;
  double tm_bnd = 0.0, tm_bulk = 0.0, tm_step = 0.0;
  int64_t tm_steps = 0;
};

namespace {

constexpr int GRAPH_W = 32;
const double SQRT_CS2_B = std::sqrt(1.0 / 3.0);

int64_t rup(int64_t v, int64_t m) { return (v + m - 1) / m * m; }

template <class V>
vl_status dupload(vl_batch* b, V*& dst, const std::vector<V>& src) {
  const size_t bytes = std::max<size_t>(src.size(), 1) * sizeof(V);
  VL_HIP(hipMalloc((void**)&dst, bytes));
  b->allocs.push_back((void*)dst);
  b->device_bytes += (int64_t)bytes;
  if (!src.empty()) VL_HIP(hipMemcpy(dst, src.data(), src.size() * sizeof(V), hipMemcpyHostToDevice));
  return VL_OK;
}

template <class V>
vl_status dalloc(vl_batch* b, V*& dst, size_t n) {
  VL_HIP(hipMalloc((void**)&dst, std::max<size_t>(n, 1) * sizeof(V)));
  b->allocs.push_back((void*)dst);
  b->device_bytes += (int64_t)(std::max<size_t>(n, 1) * sizeof(V));
  return VL_OK;
}

#define VL_TRYB(x)                \
  do {                            \
    vl_status st_ = (x);          \
    if (st_ != VL_OK) return st_; \
  } while (0)

// Host set-up of one case: the channel domain, its boundary tables, outlet nodes first.
vl_status prepare_case(const vl_batch::Pending& p, Tables& tab, vl_velocity& inlet, int32_t& npb) {
  vl_builder* b = channel_builder(&p.spec, p.body, &inlet);
  if (!b) return VL_EINVAL;
  vl_status st = build_tables(b->d, tab);
  vl_builder_free(b);
  if (st != VL_OK) return st;
  npb = outlet_nodes_first(tab);
  if (!outlet_nodes_match(tab, npb)) return set_error(VL_EINVAL, "internal: outlet-node order");
  return VL_OK;
}

// Bulk tile table for a tile shape: every (TW x TH) block of every case that holds at least one
// interior node, cases in descending step order.
void bulk_tiles(const vl_batch* b, const std::vector<int>& order, int tw, int th, std::vector<int32_t>& tc,
                std::vector<int32_t>& ty, std::vector<int32_t>& tx, std::vector<int32_t>& case_end) {
  tc.clear();
  ty.clear();
  tx.clear();
  case_end.assign(b->hcases.size(), 0);
  for (int ci : order) {
    const Tables& tab = b->tabs[ci];
    const int64_t nx = tab.nx, ny = tab.ny;
    for (int64_t y0 = 0; y0 < ny; y0 += th)
      for (int64_t x0 = 0; x0 < nx; x0 += tw) {
        bool any = false;
        for (int64_t y = y0; y < std::min(ny, y0 + th) && !any; ++y)
          for (int64_t x = x0; x < std::min(nx, x0 + tw) && !any; ++x) any = tab.flag[y * nx + x] == 0;
        if (!any) continue;
        tc.push_back(ci);
        ty.push_back((int32_t)y0);
        tx.push_back((int32_t)x0);
      }
    case_end[ci] = (int32_t)tc.size();
  }
}

constexpr int TIMING_W = 64;

// One step: the boundary kernel on the side stream and the bulk on the main stream, forked
// after the previous step and joined before the next.  `ev` (timing runs only, never inside a
// captured graph) brackets each kernel on its own stream.
template <class S>
vl_status launch_batch_step(vl_batch* b, const int64_t* tbase, int t_local, int grid_bulk,
                            const vl_batch::StepEvents* ev) {
  using T = typename S::T;
  const T* A = (const T*)b->pool[b->cur];
  T* B = (T*)b->pool[(b->cur + 1) % b->nbuf];
  // serial: the boundary kernel first, then the bulk, both on the main stream
  hipStream_t bs = b->serial ? b->stream : b->side;
  if (!b->serial) {
    VL_HIP(hipEventRecord(b->ev_fork, b->stream));
    VL_HIP(hipStreamWaitEvent(b->side, b->ev_fork, 0));
  }
  if (ev) VL_HIP(hipEventRecord(ev->bs, bs));
  hipLaunchKernelGGL(k_batch_boundary<S>, dim3((unsigned)b->grid_bnd), dim3(BT), 0, bs, A, B, b->dcases,
                     b->d_bcase, b->d_bb0, b->d_bact, tbase, t_local, b->d_kick, b->d_hist, b->d_status, b->d_done,
                     SQRT_CS2_B, (const int32_t*)nullptr, 0, (const int32_t*)nullptr, (const typename S::T*)nullptr);
  if (ev) VL_HIP(hipEventRecord(ev->be, bs));
  if (!b->serial) VL_HIP(hipEventRecord(b->ev_join, b->side));
  if (ev) VL_HIP(hipEventRecord(ev->ks, b->stream));
  hipLaunchKernelGGL(k_batch_bulk<S>, dim3((unsigned)grid_bulk), dim3(BT), 0, b->stream, A, B, b->dflags,
                     b->dcases, b->d_kcase, b->d_ky0, b->d_kx0, b->d_kact, tbase, t_local, b->tw_log2, b->d_status,
                     (const int32_t*)nullptr, 0);
  if (ev) VL_HIP(hipEventRecord(ev->ke, b->stream));
  if (!b->serial) VL_HIP(hipStreamWaitEvent(b->stream, b->ev_join, 0));
  if (ev) VL_HIP(hipEventRecord(ev->ej, b->stream));
  VL_HIP(hipGetLastError());
  b->cur = (b->cur + 1) % b->nbuf;
  b->t += 1;
  b->side0_ok = false;  // the pools' bands moved on; the level-0 side array did not
  return VL_OK;
}

// Grids of a group of K steps (workgroups): the pass, the listed tiles (or the region kernel's
// cases), the bulk tiles, and with the region kernel the boundary tiles of the cases that end
// inside the group.
struct GroupGrids {
  int pass, listed, bulk, bnd, pass_near, band, copy;
};

template <class S, int W, int V, int D, int K>
void launch_batch_pass(vl_batch* b, const vl_batch::TaskList& tl, hipStream_t st, const typename S::T* A,
                       PassOut<typename S::T> po, const int64_t* tbase, int t_local, int grid) {
  hipLaunchKernelGGL((k_batch_pass<S, W, V, D, K>), dim3((unsigned)grid), dim3(W * TB_LANES), 0, st, A, po,
                     tl.task, tl.tcase, tl.n_at, tbase, t_local, b->dcases, b->d_status);
}

// The pass kernel of the batch's K and precision: FP32 one column per lane (K = 2..5), FP16
// two (K = 2..4), FP64 one, in 4-wave workgroups (K = 2..4); 8 waves per workgroup otherwise.
template <class S>
vl_status launch_pass_k(vl_batch* b, const vl_batch::TaskList& tl, hipStream_t st, const typename S::T* A,
                        PassOut<typename S::T> po, const int64_t* tbase, int t_local, int grid) {
  const int K = b->pass_k;
  if constexpr (std::is_same<S, StoreF16>::value) {
    if (K == 2) launch_batch_pass<S, 8, 2, 2, 2>(b, tl, st, A, po, tbase, t_local, grid);
    else if (K == 3) launch_batch_pass<S, 8, 2, 2, 3>(b, tl, st, A, po, tbase, t_local, grid);
    else if (K == 4) launch_batch_pass<S, 8, 2, 2, 4>(b, tl, st, A, po, tbase, t_local, grid);
    else return set_error(VL_EINVAL, "FP16 batch passes take 2..4 steps");
  } else if constexpr (std::is_same<S, StoreF64>::value) {
    if (K == 2) launch_batch_pass<S, 4, 1, 3, 2>(b, tl, st, A, po, tbase, t_local, grid);
    else if (K == 3) launch_batch_pass<S, 4, 1, 3, 3>(b, tl, st, A, po, tbase, t_local, grid);
    else if (K == 4) launch_batch_pass<S, 4, 1, 3, 4>(b, tl, st, A, po, tbase, t_local, grid);
    else return set_error(VL_EINVAL, "FP64 batch passes take 2..4 steps");
  } else {
    if (K == 2) launch_batch_pass<S, 8, 1, 3, 2>(b, tl, st, A, po, tbase, t_local, grid);
    else if (K == 3) launch_batch_pass<S, 8, 1, 3, 3>(b, tl, st, A, po, tbase, t_local, grid);
    else if (K == 4) launch_batch_pass<S, 8, 1, 3, 4>(b, tl, st, A, po, tbase, t_local, grid);
    else if (K == 5) launch_batch_pass<S, 8, 1, 3, 5>(b, tl, st, A, po, tbase, t_local, grid);
    else return set_error(VL_EINVAL, "FP32 batch passes take 2..5 steps");
  }
  return VL_OK;
}

// The single-step part of step j of a group (reads pool cur + j, writes pool cur + j + 1): the
// boundary kernel on stream `bs`, the listed nodes and the bulk tiles of the cases that end inside
// the group on stream `ms`, the band nodes on stream `ts`.
template <class S>
void launch_group_phase(vl_batch* b, hipStream_t bs, hipStream_t ms, hipStream_t ts, const typename S::T* A,
                        typename S::T* B, const int64_t* tbase, int t_local, int j, const GroupGrids& gg) {
  using T = typename S::T;
  const int K = b->pass_k;
  // the bands' time t + j: at the group's first step the level-0 side array (the pool when a
  // single step ran since the last group), then the side array of level j; time t + j + 1: the
  // side array of level j + 1, at the group's last step the level-0 side array, copied into
  // the pool after the band kernel
  const T* side_rd = !b->band ? nullptr
                     : j > 0  ? (const T*)b->bside[j - 1]
                              : (b->side0_ok ? (const T*)b->bside0 : nullptr);
  T* side_wr = !b->band ? nullptr : j + 1 < K ? (T*)b->bside[j] : (T*)b->bside0;
  hipLaunchKernelGGL(k_batch_boundary<S>, dim3((unsigned)b->grid_bnd), dim3(BT), 0, bs, A, B, b->dcases, b->d_bcase,
                     b->d_bb0, b->d_bact, tbase, t_local + j, b->d_kick, b->d_hist, b->d_status, b->d_done,
                     SQRT_CS2_B, (const int32_t*)nullptr, j, b->band ? (const int32_t*)b->d_bfirst_at : nullptr,
                     side_rd);
  hipLaunchKernelGGL(k_batch_listed<S>, dim3((unsigned)gg.listed), dim3(BT), 0, ms, A, B, b->dcases, b->d_lcase,
                     b->d_lfirst, b->d_lcount, b->d_lnodes, b->d_ln_at, tbase, t_local + j, j, b->d_status, side_rd);
  if (b->band) {
    hipLaunchKernelGGL(k_batch_band<S>, dim3((unsigned)gg.band), dim3(BT), 0, ts, A, B, side_rd, side_wr, b->dcases,
                       b->d_btile, b->d_nband_at, b->d_bidx, tbase, t_local + j, j, b->d_kick, b->d_status,
                       SQRT_CS2_B);
    if (j + 1 == K)
      hipLaunchKernelGGL(k_band_to_pool<T>, dim3((unsigned)gg.copy), dim3(BT), 0, ts, B, (const T*)b->bside0,
                         b->dcases, b->d_ctile, b->d_ncopy_at, b->d_bidx, tbase, t_local + j, j, b->d_status);
  }
  hipLaunchKernelGGL(k_batch_bulk<S>, dim3((unsigned)gg.bulk), dim3(BT), 0, ms, A, B, b->dflags, b->dcases,
                     b->d_kcase, b->d_ky0, b->d_kx0, b->d_kact, tbase, t_local + j, b->tw_log2, b->d_status,
                     (const int32_t*)b->d_kfirst_at, j);
}

// A group of K steps from t = *tbase + t_local, pools cur .. cur + K.  The passes take the
// covered interior from t to t + K (pool cur to pool cur + K; tasks with ring nodes also write
// their levels 1 .. K-1 into pools cur + 1 .. cur + K - 1); step j of the single-step part reads
// pool cur + j and writes pool cur + j + 1, after the passes, so that its boundary and listed
// nodes find their covered neighbours' level j written.
//  * Step 0: the pass (main stream) beside step 0's single-step part (side stream).  The pass
//    reads pool cur and writes covered nodes only; step 0's boundary and listed nodes read pool
//    cur and write pool cur + 1.
//  * Steps 1 .. K-1: the boundary kernel on the side stream beside the listed and bulk kernels.
// Every step joins before the next.  (Measured alternatives, all slower on the 512-case sweep:
// overlapping group g's single steps with group g + 1's passes of the tasks that depend on
// covered nodes only, as the single-domain path does, since the batch's single-step kernels
// slow the passes more than they hide; and one pass kernel per collision model, side by side
// or one after the other, since each case's tasks then no longer run together.)
// The region kernel over steps j0 .. j1 - 1 of the group.
template <class S, int RT>
void launch_region_rt(vl_batch* b, hipStream_t st, const PoolSet<typename S::T>& ps, const int64_t* tbase,
                      int t_local, int j0, int j1, int grid) {
  hipLaunchKernelGGL((k_batch_region<S, RT>), dim3((unsigned)grid), dim3(RT), 0, st, ps, b->dcases, b->d_rcase,
                     b->d_rn_at, b->d_clfirst, b->d_clcount, b->d_lnodes, tbase, t_local, j0, j1, b->d_kick,
                     b->d_hist, b->d_status, b->d_done, SQRT_CS2_B);
}
template <class S>
void launch_region(vl_batch* b, hipStream_t st, const PoolSet<typename S::T>& ps, const int64_t* tbase, int t_local,
                   int j0, int j1, int grid) {
  if (b->region_threads == 256) launch_region_rt<S, 256>(b, st, ps, tbase, t_local, j0, j1, grid);
  else if (b->region_threads == 1024) launch_region_rt<S, 1024>(b, st, ps, tbase, t_local, j0, j1, grid);
  else launch_region_rt<S, 512>(b, st, ps, tbase, t_local, j0, j1, grid);
}

// Step j of the cases that end inside the group (single steps of all their nodes; the region
// kernel and the pass take the other cases): the boundary tiles past bfirst_at[t], then the bulk
// tiles past kfirst_at[t].
template <class S>
void launch_finishing_step(vl_batch* b, hipStream_t st, const typename S::T* A, typename S::T* B,
                           const int64_t* tbase, int t_local, int j, const GroupGrids& gg) {
  hipLaunchKernelGGL(k_batch_boundary<S>, dim3((unsigned)gg.bnd), dim3(BT), 0, st, A, B, b->dcases, b->d_bcase,
                     b->d_bb0, b->d_bact, tbase, t_local + j, b->d_kick, b->d_hist, b->d_status, b->d_done,
                     SQRT_CS2_B, (const int32_t*)b->d_bfirst_at, j, (const int32_t*)nullptr,
                     (const typename S::T*)nullptr);
  hipLaunchKernelGGL(k_batch_bulk<S>, dim3((unsigned)gg.bulk), dim3(BT), 0, st, A, B, b->dflags, b->dcases,
                     b->d_kcase, b->d_ky0, b->d_kx0, b->d_kact, tbase, t_local + j, b->tw_log2, b->d_status,
                     (const int32_t*)b->d_kfirst_at, j);
}

// A group of K steps with the region kernel.  Main stream: the pass, then the region kernel's
// steps 1 .. K-1 (or 0 .. K-1).  Side stream: the region kernel's step 0 beside the pass (it reads
// pool cur only), then the K single steps of the cases that end inside the group.  Joined at the
// end.
template <class S>
vl_status launch_batch_group_region(vl_batch* b, const int64_t* tbase, int t_local, const GroupGrids& gg) {
  using T = typename S::T;
  const int K = b->pass_k, nb = b->nbuf;
  auto P = [&](int j) { return (T*)b->pool[(b->cur + j) % nb]; };
  PassOut<T> po{};
  for (int L = 1; L < K; ++L) po.ring[L - 1] = P(L);
  po.out = P(K);
  PoolSet<T> ps{};
  for (int j = 0; j <= K; ++j) ps.p[j] = P(j);
  if (b->serial) {
    VL_TRYB((launch_pass_k<S>(b, b->tasks, b->stream, P(0), po, tbase, t_local, gg.pass)));
    launch_region<S>(b, b->stream, ps, tbase, t_local, 0, K, gg.listed);
    for (int j = 0; j < K; ++j) launch_finishing_step<S>(b, b->stream, P(j), P(j + 1), tbase, t_local, j, gg);
    VL_HIP(hipGetLastError());
  } else {
    VL_HIP(hipEventRecord(b->ev_fork, b->stream));
    VL_HIP(hipStreamWaitEvent(b->side, b->ev_fork, 0));
    VL_TRYB((launch_pass_k<S>(b, b->tasks, b->stream, P(0), po, tbase, t_local, gg.pass)));
    if (b->region0_side) {
      launch_region<S>(b, b->side, ps, tbase, t_local, 0, 1, gg.listed);
      VL_HIP(hipEventRecord(b->ev_mid, b->side));
    }
    for (int j = 0; j < K; ++j) launch_finishing_step<S>(b, b->side, P(j), P(j + 1), tbase, t_local, j, gg);
    VL_HIP(hipEventRecord(b->ev_join, b->side));
    if (b->region0_side) VL_HIP(hipStreamWaitEvent(b->stream, b->ev_mid, 0));
    launch_region<S>(b, b->stream, ps, tbase, t_local, b->region0_side ? 1 : 0, K, gg.listed);
    VL_HIP(hipStreamWaitEvent(b->stream, b->ev_join, 0));
    VL_HIP(hipGetLastError());
  }
  b->cur = (b->cur + K) % nb;
  b->t += K;
  return VL_OK;
}

// A group of K steps with the region kernel beside the next group's far pass.
//  * Main stream: the far tasks' pass, then (after the previous group's region kernel) the near
//    tasks' pass.  A far task's outputs lie more than K nodes from every uncovered node, so its
//    pass reads covered nodes only (of the previous group's level K, written by both passes
//    before it in stream order) and writes neither ring values nor anything the region kernel
//    touches; a near task's region holds uncovered nodes, whose level K the previous group's
//    region kernel writes.
//  * Side stream: the single steps of the cases that end inside the group, then (after the near
//    pass, which writes every ring value) the region kernel's K steps.
// So the region kernel of group g runs beside the far pass of group g + 1.  Each event is
// waited for before it is recorded again; join_side joins the side stream back at the end of a
// window or a run.
template <class S>
vl_status launch_batch_group_overlap(vl_batch* b, const int64_t* tbase, int t_local, const GroupGrids& gg) {
  using T = typename S::T;
  const int K = b->pass_k, nb = b->nbuf;
  auto P = [&](int j) { return (T*)b->pool[(b->cur + j) % nb]; };
  PassOut<T> po{};
  for (int L = 1; L < K; ++L) po.ring[L - 1] = P(L);
  po.out = P(K);
  PoolSet<T> ps{};
  for (int j = 0; j <= K; ++j) ps.p[j] = P(j);
  bool first = false;
  if (!b->ov_open) {
    VL_HIP(hipEventRecord(b->ev_fork, b->stream));
    VL_HIP(hipStreamWaitEvent(b->side, b->ev_fork, 0));
    b->ov_open = true;
    first = true;
  }
  VL_TRYB((launch_pass_k<S>(b, b->tasks_far, b->stream, P(0), po, tbase, t_local, gg.pass)));
  if (!first) VL_HIP(hipStreamWaitEvent(b->stream, b->ev_join, 0));
  VL_TRYB((launch_pass_k<S>(b, b->tasks_near, b->stream, P(0), po, tbase, t_local, gg.pass_near)));
  VL_HIP(hipEventRecord(b->ev_mid, b->stream));
  for (int j = 0; j < K; ++j) launch_finishing_step<S>(b, b->side, P(j), P(j + 1), tbase, t_local, j, gg);
  VL_HIP(hipStreamWaitEvent(b->side, b->ev_mid, 0));
  launch_region<S>(b, b->side, ps, tbase, t_local, 0, K, gg.listed);
  VL_HIP(hipEventRecord(b->ev_join, b->side));
  VL_HIP(hipGetLastError());
  b->cur = (b->cur + K) % nb;
  b->t += K;
  return VL_OK;
}

// Joins the side stream back into the main stream after overlapped groups.
vl_status join_side(vl_batch* b) {
  if (!b->ov_open) return VL_OK;
  VL_HIP(hipStreamWaitEvent(b->stream, b->ev_join, 0));
  b->ov_open = false;
  return VL_OK;
}

// A group of K steps whose single steps 1 .. K-1 run beside the next group's far pass.
//  * Main stream: the far tasks' pass, then (after the previous group's single steps) the near
//    tasks' pass.  A far task's outputs lie more than K nodes from every uncovered node, so it
//    reads covered nodes only (level K of the previous group, written by both passes before it
//    in stream order) and writes no ring value; its level-K output lands in the pool that the
//    previous group's last single steps still use, but only at far nodes, which no single step
//    reads or writes.
//  * Side stream: step 0's single-step part, in stream order after the previous group's single
//    steps and beside this group's passes (it reads pool cur only and writes uncovered nodes and
//    side arrays, which the passes do not touch), then, after the near pass (both passes have
//    then written every ring value), steps 1 .. K-1: the boundary kernel on the side stream, the
//    band kernel (and the bands' copy into the pool after the last step) on the third stream,
//    the listed and bulk kernels on the fourth, forked from and joined back into the side
//    stream per step.  The near pass of the next group waits for the end of these steps.
// join_side joins the side stream back at the end of a window or a run.
template <class S>
vl_status launch_batch_group_steps_overlap(vl_batch* b, const int64_t* tbase, int t_local, const GroupGrids& gg) {
  using T = typename S::T;
  const int K = b->pass_k, nb = b->nbuf;
  auto P = [&](int j) { return (T*)b->pool[(b->cur + j) % nb]; };
  PassOut<T> po{};
  for (int L = 1; L < K; ++L) po.ring[L - 1] = P(L);
  po.out = P(K);
  bool first = false;
  if (!b->ov_open) {
    VL_HIP(hipEventRecord(b->ev_fork, b->stream));
    VL_HIP(hipStreamWaitEvent(b->side, b->ev_fork, 0));
    b->ov_open = true;
    first = true;
  }
  VL_TRYB((launch_pass_k<S>(b, b->tasks_far, b->stream, P(0), po, tbase, t_local, gg.pass)));
  if (!first) VL_HIP(hipStreamWaitEvent(b->stream, b->ev_join, 0));
  VL_TRYB((launch_pass_k<S>(b, b->tasks_near, b->stream, P(0), po, tbase, t_local, gg.pass_near)));
  VL_HIP(hipEventRecord(b->ev_mid, b->stream));
  launch_group_phase<S>(b, b->side, b->side, b->side, P(0), P(1), tbase, t_local, 0, gg);
  VL_HIP(hipStreamWaitEvent(b->side, b->ev_mid, 0));
  for (int j = 1; j < K; ++j) {
    VL_HIP(hipEventRecord(b->ev_sfork, b->side));
    VL_HIP(hipStreamWaitEvent(b->third, b->ev_sfork, 0));
    VL_HIP(hipStreamWaitEvent(b->fourth, b->ev_sfork, 0));
    launch_group_phase<S>(b, b->side, b->fourth, b->third, P(j), P(j + 1), tbase, t_local, j, gg);
    VL_HIP(hipEventRecord(b->ev_join3, b->third));
    VL_HIP(hipEventRecord(b->ev_join4, b->fourth));
    VL_HIP(hipStreamWaitEvent(b->side, b->ev_join3, 0));
    VL_HIP(hipStreamWaitEvent(b->side, b->ev_join4, 0));
  }
  VL_HIP(hipEventRecord(b->ev_join, b->side));
  VL_HIP(hipGetLastError());
  b->cur = (b->cur + K) % nb;
  b->t += K;
  // the last step wrote the bands' level K into the level-0 side array as well
  b->side0_ok = b->band;
  return VL_OK;
}

template <class S>
vl_status launch_batch_group(vl_batch* b, const int64_t* tbase, int t_local, const GroupGrids& gg) {
  using T = typename S::T;
  if (b->region && b->overlap && !b->serial) return launch_batch_group_overlap<S>(b, tbase, t_local, gg);
  if (!b->region && b->step_overlap) return launch_batch_group_steps_overlap<S>(b, tbase, t_local, gg);
  if (b->region) return launch_batch_group_region<S>(b, tbase, t_local, gg);
  const int K = b->pass_k, nb = b->nbuf;
  auto P = [&](int j) { return (T*)b->pool[(b->cur + j) % nb]; };
  PassOut<T> po{};
  for (int L = 1; L < K; ++L) po.ring[L - 1] = P(L);
  po.out = P(K);
  if (b->serial) {
    VL_TRYB((launch_pass_k<S>(b, b->tasks, b->stream, P(0), po, tbase, t_local, gg.pass)));
    for (int j = 0; j < K; ++j)
      launch_group_phase<S>(b, b->stream, b->stream, b->stream, P(j), P(j + 1), tbase, t_local, j, gg);
    VL_HIP(hipGetLastError());
  } else {
    VL_HIP(hipEventRecord(b->ev_fork, b->stream));
    VL_HIP(hipStreamWaitEvent(b->side, b->ev_fork, 0));
    VL_TRYB((launch_pass_k<S>(b, b->tasks, b->stream, P(0), po, tbase, t_local, gg.pass)));
    launch_group_phase<S>(b, b->side, b->side, b->side, P(0), P(1), tbase, t_local, 0, gg);
    VL_HIP(hipEventRecord(b->ev_join, b->side));
    VL_HIP(hipStreamWaitEvent(b->stream, b->ev_join, 0));
    // steps 1 .. K-1: the band kernel on a third stream when there are bands, so that the band,
    // listed and boundary nodes, each bound by latency, run side by side
    const bool third = b->band && b->third;
    for (int j = 1; j < K; ++j) {
      VL_HIP(hipEventRecord(b->ev_fork, b->stream));
      VL_HIP(hipStreamWaitEvent(b->side, b->ev_fork, 0));
      if (third) VL_HIP(hipStreamWaitEvent(b->third, b->ev_fork, 0));
      launch_group_phase<S>(b, b->side, b->stream, third ? b->third : b->stream, P(j), P(j + 1), tbase, t_local, j,
                            gg);
      VL_HIP(hipEventRecord(b->ev_join, b->side));
      if (third) VL_HIP(hipEventRecord(b->ev_join3, b->third));
      VL_HIP(hipStreamWaitEvent(b->stream, b->ev_join, 0));
      if (third) VL_HIP(hipStreamWaitEvent(b->stream, b->ev_join3, 0));
    }
    VL_HIP(hipGetLastError());
  }
  b->cur = (b->cur + K) % nb;
  b->t += K;
  // the last step wrote the bands' level K into the level-0 side array as well
  b->side0_ok = b->band;
  return VL_OK;
}

vl_status batch_group(vl_batch* b, const int64_t* tbase, int t_local, const GroupGrids& gg) {
  switch (b->precision) {
    case VL_F64: return launch_batch_group<StoreF64>(b, tbase, t_local, gg);
    case VL_F16: return launch_batch_group<StoreF16>(b, tbase, t_local, gg);
    default: return launch_batch_group<StoreF32>(b, tbase, t_local, gg);
  }
}

vl_status batch_step(vl_batch* b, const int64_t* tbase, int t_local, int grid_bulk,
                     const vl_batch::StepEvents* ev = nullptr) {
  switch (b->precision) {
    case VL_F64: return launch_batch_step<StoreF64>(b, tbase, t_local, grid_bulk, ev);
    case VL_F16: return launch_batch_step<StoreF16>(b, tbase, t_local, grid_bulk, ev);
    default: return launch_batch_step<StoreF32>(b, tbase, t_local, grid_bulk, ev);
  }
}

// Geometric ladder of grid sizes (each about 9 % above the last, the last one = all tiles).
void make_ladder(vl_batch* b) {
  b->ladder.clear();
  int64_t v = 1;
  while (v < b->nbulk_tiles) {
    b->ladder.push_back((int32_t)v);
    v = std::max(v + 1, (v * 109 + 99) / 100);
  }
  b->ladder.push_back(std::max(1, b->nbulk_tiles));
}

// Bulk grid for a span of steps starting at global step t: the active-tile count can only
// shrink with t, so the count at t covers the span.  The kernel loops over tiles when the
// grid is smaller than the count, so any grid is correct; this one only sets the speed.
int bulk_grid_at(const vl_batch* b, int64_t t) {
  if (b->bulk_mult > 0) return b->grid_bulk;
  const size_t i = (size_t)std::min<int64_t>(std::max<int64_t>(t, 0), (int64_t)b->h_kact.size() - 1);
  const int32_t need = std::max<int32_t>(b->h_kact[i], 1);
  auto it = std::lower_bound(b->ladder.begin(), b->ladder.end(), need);
  return it == b->ladder.end() ? b->ladder.back() : *it;
}

// The smallest value of the grid ladder (1, then about 9 % steps) that is >= need, capped at all.
int ladder_ceil(int64_t need, int64_t all) {
  all = std::max<int64_t>(all, 1);
  int64_t v = 1;
  while (v < need && v < all) v = std::max(v + 1, (v * 109 + 99) / 100);
  return (int)std::min(v, all);
}

// Grids of the groups of K steps starting at t, t + K, ..., t + (groups - 1) K.  The pass must
// cover every task it may run (its count only shrinks with t, so the count at t covers the
// span); the listed and bulk kernels loop over their tiles, so their grids only set the speed.
GroupGrids group_grids_at(const vl_batch* b, int64_t t, int groups) {
  const int K = b->pass_k, W = b->pass_waves;
  const int64_t last = (int64_t)b->h_ln_at.size() - 1;
  auto at = [&](const std::vector<int32_t>& v, int64_t i) { return v[(size_t)std::min(std::max<int64_t>(i, 0), last)]; };
  GroupGrids gg;
  // far and near tasks in launches of their own
  const bool split = (b->region && b->overlap && !b->serial) || (!b->region && b->step_overlap);
  const vl_batch::TaskList& first_list = split ? b->tasks_far : b->tasks;
  gg.pass = ladder_ceil((at(first_list.h_n_at, t) + W - 1) / W, (first_list.n + W - 1) / W);
  gg.pass_near = split ? ladder_ceil((at(b->tasks_near.h_n_at, t) + W - 1) / W, (b->tasks_near.n + W - 1) / W) : 0;
  gg.listed = b->region ? ladder_ceil(at(b->h_rn_at, t), (int64_t)b->hcases.size())
                        : ladder_ceil(at(b->h_ln_at, t), b->nlisted_tiles);
  int64_t bulk = 1, bnd = 1;
  for (int g = 0; g < groups; ++g)
    for (int j = 0; j < K; ++j) {
      const int64_t tg = t + (int64_t)g * K;
      bulk = std::max<int64_t>(bulk, (int64_t)at(b->h_kact, tg + j) - at(b->h_kfirst_at, tg));
      if (b->region) bnd = std::max<int64_t>(bnd, (int64_t)at(b->h_bact, tg + j) - at(b->h_bfirst_at, tg));
    }
  gg.bulk = ladder_ceil(bulk, b->nbulk_tiles);
  gg.bnd = b->region ? ladder_ceil(bnd, b->nbnd_tiles) : 0;
  gg.band = b->band ? ladder_ceil(at(b->h_nband_at, t), b->nband_tiles) : 0;
  gg.copy = b->band ? ladder_ceil(at(b->h_ncopy_at, t), b->ncopy_tiles) : 0;
  return gg;
}

// Adds the kernel and step times of the first n steps' events to the running totals.
vl_status read_timing(vl_batch* b, int n) {
  VL_HIP(hipStreamSynchronize(b->stream));
  for (int i = 0; i < n; ++i) {
    const auto& e = b->tev[(size_t)i];
    float bnd = 0.0f, bulk = 0.0f, step = 0.0f;
    VL_HIP(hipEventElapsedTime(&bnd, e.bs, e.be));
    VL_HIP(hipEventElapsedTime(&bulk, e.ks, e.ke));
    VL_HIP(hipEventElapsedTime(&step, e.bs, e.ej));
    b->tm_bnd += bnd;
    b->tm_bulk += bulk;
    b->tm_step += step;
    b->tm_steps += 1;
  }
  return VL_OK;
}

// A trial of step 0 for the autotuner: the boundary kernel and a bulk pass with candidate tile
// tables, arranged as a real step (side by side, or one after the other in serial mode).  Step 0
// reads only A and the primed outlet velocities, and writes B, the outlet offsets and the
// step-0 force terms and velocities: the same values every time, so trials change nothing a
// real step 0 would not write again.
template <class S>
vl_status launch_trial_step(vl_batch* b, int grid, int tw_log2, const int32_t* kc, const int32_t* ky,
                            const int32_t* kx, const int32_t* kact) {
  using T = typename S::T;
  const T* A = (const T*)b->pool[b->cur];
  T* B = (T*)b->pool[(b->cur + 1) % b->nbuf];
  const int64_t* t0 = b->d_tbase + 1;  // a zero
  hipStream_t bs = b->serial ? b->stream : b->side;
  if (!b->serial) {
    VL_HIP(hipEventRecord(b->ev_fork, b->stream));
    VL_HIP(hipStreamWaitEvent(b->side, b->ev_fork, 0));
  }
  hipLaunchKernelGGL(k_batch_boundary<S>, dim3((unsigned)b->grid_bnd), dim3(BT), 0, bs, A, B, b->dcases, b->d_bcase,
                     b->d_bb0, b->d_bact, t0, 0, b->d_kick, b->d_hist, b->d_status, b->d_done, SQRT_CS2_B,
                     (const int32_t*)nullptr, 0, (const int32_t*)nullptr, (const typename S::T*)nullptr);
  if (!b->serial) VL_HIP(hipEventRecord(b->ev_join, b->side));
  hipLaunchKernelGGL(k_batch_bulk<S>, dim3((unsigned)grid), dim3(BT), 0, b->stream, A, B, b->dflags, b->dcases, kc,
                     ky, kx, kact, t0, 0, tw_log2, b->d_status, (const int32_t*)nullptr, 0);
  if (!b->serial) VL_HIP(hipStreamWaitEvent(b->stream, b->ev_join, 0));
  VL_HIP(hipGetLastError());
  return VL_OK;
}

vl_status trial_step(vl_batch* b, int grid, int tw_log2, const int32_t* kc, const int32_t* ky, const int32_t* kx,
                     const int32_t* kact) {
  switch (b->precision) {
    case VL_F64: return launch_trial_step<StoreF64>(b, grid, tw_log2, kc, ky, kx, kact);
    case VL_F16: return launch_trial_step<StoreF16>(b, grid, tw_log2, kc, ky, kx, kact);
    default: return launch_trial_step<StoreF32>(b, grid, tw_log2, kc, ky, kx, kact);
  }
}

void free_batch(vl_batch* b) {
  if (!b) return;
  if (b->stream) (void)hipStreamSynchronize(b->stream);
  if (b->side) (void)hipStreamSynchronize(b->side);
  for (auto& p : b->pending)
    if (p.body) vl_shape_free(p.body);
  for (void* p : b->allocs) (void)hipFree(p);
  for (void* p : b->pool)
    if (p) (void)hipFree(p);
  for (void* p : b->bside)
    if (p) (void)hipFree(p);
  if (b->bside0) (void)hipFree(b->bside0);
  for (auto& kv : b->graphs)
    if (kv.second) (void)hipGraphExecDestroy(kv.second);
  for (auto& e : b->tev)
    for (hipEvent_t h : {e.bs, e.be, e.ks, e.ke, e.ej})
      if (h) (void)hipEventDestroy(h);
  if (b->ev_fork) (void)hipEventDestroy(b->ev_fork);
  if (b->ev_join) (void)hipEventDestroy(b->ev_join);
  if (b->ev_mid) (void)hipEventDestroy(b->ev_mid);
  if (b->ev_join3) (void)hipEventDestroy(b->ev_join3);
  if (b->ev_sfork) (void)hipEventDestroy(b->ev_sfork);
  if (b->ev_join4) (void)hipEventDestroy(b->ev_join4);
  if (b->third) (void)hipStreamDestroy(b->third);
  if (b->fourth) (void)hipStreamDestroy(b->fourth);
  if (b->side) (void)hipStreamDestroy(b->side);
  if (b->stream) (void)hipStreamDestroy(b->stream);
  delete b;
}

vl_status build(vl_batch* b) {
  const size_t nc = b->pending.size();
  if (nc == 0) return set_error(VL_EINVAL, "batch has no cases");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return set_error(VL_EDEVICE, "no HIP device");
  using clk = std::chrono::steady_clock;
  auto t_mark = clk::now();
  auto lap = [&](int k) {
    const auto now = clk::now();
    b->setup_ms[k] += std::chrono::duration<double, std::milli>(now - t_mark).count();
    t_mark = now;
  };
  VL_HIP(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking));
  VL_HIP(hipStreamCreateWithFlags(&b->side, hipStreamNonBlocking));
  VL_HIP(hipEventCreateWithFlags(&b->ev_fork, hipEventDisableTiming));
  VL_HIP(hipEventCreateWithFlags(&b->ev_join, hipEventDisableTiming));
  VL_HIP(hipEventCreateWithFlags(&b->ev_mid, hipEventDisableTiming));
  // the band kernel's own stream (VL_BATCH_THIRD=0: the band kernel on the main stream)
  if (const char* e = std::getenv("VL_BATCH_THIRD"); !e || e[0] != '0') {
    VL_HIP(hipStreamCreateWithFlags(&b->third, hipStreamNonBlocking));
    VL_HIP(hipEventCreateWithFlags(&b->ev_join3, hipEventDisableTiming));
    VL_HIP(hipStreamCreateWithFlags(&b->fourth, hipStreamNonBlocking));
    VL_HIP(hipEventCreateWithFlags(&b->ev_sfork, hipEventDisableTiming));
    VL_HIP(hipEventCreateWithFlags(&b->ev_join4, hipEventDisableTiming));
  }
  // ---- set-up of every case: on the device when every body is a polygon or a circle (the set-up
  // kernels evaluate those predicates themselves), else on host threads; VL_BATCH_SETUP=host
  // forces the host builder ----
  b->tabs.assign(nc, Tables{});
  std::vector<vl_velocity> inlet(nc);
  std::vector<int32_t> npb(nc, 0);
  std::vector<vl_status> st(nc, VL_OK);
  std::vector<std::string> err(nc);
  std::vector<ShapeDesc> desc(nc);
  b->setup_on_device = true;
  if (const char* e = std::getenv("VL_BATCH_SETUP")) b->setup_on_device = std::string(e) != "host";
  for (size_t i = 0; i < nc && b->setup_on_device; ++i) {
    desc[i] = b->pending[i].body->s->describe();
    b->setup_on_device = desc[i].kind != ShapeDesc::NONE;
  }
  if (b->setup_on_device) {
    for (size_t i = 0; i < nc; ++i) {
      b->tabs[i].nx = b->pending[i].spec.nx;
      b->tabs[i].ny = b->pending[i].spec.ny;
    }
  } else {
    const unsigned nth = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    std::vector<std::thread> pool;
    for (unsigned w = 0; w < nth; ++w)
      pool.emplace_back([&, w] {
        for (size_t i = w; i < nc; i += nth) {
          st[i] = prepare_case(b->pending[i], b->tabs[i], inlet[i], npb[i]);
          b->pending[i].body = nullptr;  // consumed by the builder
          if (st[i] != VL_OK) err[i] = vl_last_error();
        }
      });
    for (auto& th : pool) th.join();
  }
  for (size_t i = 0; i < nc; ++i)
    if (st[i] != VL_OK) return set_error(st[i], "case " + std::to_string(i) + ": " + err[i]);
  // ---- packing ----
  b->hcases.assign(nc, BatchCase{});
  int64_t pop_elems = 0, flag_bytes = 0;
  std::vector<double> kick_tab;
  int64_t hist_len = 0;
  for (size_t i = 0; i < nc; ++i) {
    const auto& p = b->pending[i];
    const Tables& tab = b->tabs[i];
    BatchCase& c = b->hcases[i];
    c.g.nx = (int32_t)tab.nx;
    c.g.ny = (int32_t)tab.ny;
    c.g.pitch = (int32_t)rup(tab.nx, 64);
    c.g.periodic_x = c.g.periodic_y = 0;
    c.g.plane = rup((int64_t)c.g.pitch * tab.ny, 64);
    double op, om;
    const double tau = p.spec.nu / (1.0 / 3.0) + 0.5;
    omegas(p.spec.collision, p.spec.magic, tau, op, om);
    c.c = Coll{op, om, 0.0, 0.0, (float)op, (float)om, 0.0f, 0.0f, p.spec.collision == VL_BGK ? MODEL_BGK : MODEL_TRT};
    c.pop_off = pop_elems;
    pop_elems += Q * c.g.plane;
    c.flag_off = flag_bytes;
    flag_bytes += (int64_t)c.g.pitch * tab.ny;
    c.u_ref = p.spec.u;
    c.coef_k = 2.0 / (p.spec.u * p.spec.u * p.d);  // Channel::coefficients
    c.ramp_steps = p.spec.ramp_steps;
    c.kick_start = p.spec.has_kick ? (int64_t)p.spec.kick.start : 0;
    c.kick_dur = p.spec.has_kick ? (int64_t)p.spec.kick.duration : 0;
    c.kick_off = (int64_t)kick_tab.size();
    for (int64_t k = 0; k < c.kick_dur; ++k)
      kick_tab.push_back(channel_kick(p.spec.kick, true, (uint64_t)(c.kick_start + k)));
    c.steps = p.steps;
    c.hist_off = hist_len;
    hist_len += 2 * p.steps;
    c.absorbing = 1;
    c.inflow = p.spec.inflow;
    c.inflow_u = p.spec.u;
    c.inflow_h = (double)(tab.ny - 2);
    b->total_nodes += (int64_t)tab.nx * tab.ny;
    b->max_steps = std::max(b->max_steps, p.steps);
    (void)inlet;
  }
  // ---- device set-up: flags straight into the batch's pitched flag array, tables into packed
  // pools at the host builder's offsets; the flags come back for the pass planner ----
  SetupTables dtab{};
  std::vector<SetupCounts> dcount;
  if (b->setup_on_device) {
    std::vector<SetupCase> sc(nc);
    std::vector<double> verts;
    for (size_t i = 0; i < nc; ++i) {
      const auto& sp = b->pending[i].spec;
      const BatchCase& c = b->hcases[i];
      SetupCase& k = sc[i];
      k = SetupCase{};
      k.nx = c.g.nx;
      k.ny = c.g.ny;
      k.pitch = c.g.pitch;
      k.inflow = sp.inflow;
      k.cell0 = c.flag_off;
      k.u = sp.u;
      k.h = (double)(sp.ny - 2);
      k.has_kick = sp.has_kick != 0;
      k.kcx = sp.kick.cx;
      k.kcy = sp.kick.cy;
      k.body = desc[i].kind;
      k.bcx = desc[i].cx;
      k.bcy = desc[i].cy;
      k.br = desc[i].r;
      k.vert0 = (int64_t)verts.size() / 2;
      k.nverts = (int32_t)desc[i].pts.size();
      for (const auto& v : desc[i].pts) {
        verts.push_back(v[0]);
        verts.push_back(v[1]);
      }
    }
    VL_TRYB(dalloc(b, b->dflags, (size_t)flag_bytes));
    const SetupAlloc alloc = [b](void** p, size_t bytes) {
      VL_HIP(hipMalloc(p, bytes));
      b->allocs.push_back(*p);
      b->device_bytes += (int64_t)bytes;
      return VL_OK;
    };
    VL_TRYB(device_setup(sc, verts, b->dflags, flag_bytes, b->stream, alloc, dcount, dtab));
    std::vector<uint8_t> fl((size_t)flag_bytes);
    VL_HIP(hipMemcpy(fl.data(), b->dflags, (size_t)flag_bytes, hipMemcpyDeviceToHost));
    for (size_t i = 0; i < nc; ++i) {
      Tables& tab = b->tabs[i];
      const BatchCase& c = b->hcases[i];
      tab.fluid_nodes = dcount[i].fluid;
      tab.flag.resize((size_t)(tab.nx * tab.ny));
      for (int64_t y = 0; y < tab.ny; ++y)
        std::memcpy(&tab.flag[(size_t)(y * tab.nx)], &fl[(size_t)(c.flag_off + y * c.g.pitch)], (size_t)tab.nx);
      npb[i] = (int32_t)dcount[i].npb;
    }
  }
  for (size_t i = 0; i < nc; ++i) b->total_fluid += b->tabs[i].fluid_nodes;
  lap(0);
  // ---- K-step passes: the plan of every case (its strips' tasks and listed nodes) ----
  // Defaults: FP32 5 steps per pass (the 512-case sweep on MI355X: 93 GLUPS against 87 at 4 and
  // 67 in single steps), FP64 3 (the 4-step FP64 pass of both collides needs all 256 registers),
  // FP16 3; tasks of 14 (FP32), 16 (FP64) or 24 (FP16) rows.  FP32 on the sweep, cases/s by task
  // height with the bands in side arrays: 10 rows 32.8, 12 32.9, 14 33.2, 16 32.8, 20 32.3
  // (before them: 8 rows 25.9, 10 26.1, 12 26.3, 14 26.2, 16 25.9, 20 25.4, 28 24.4, 36 23.7);
  // shorter tasks re-read more halo rows but balance better over the sweep's short, narrow cases.
  // VL_BATCH_PASS_STEPS (1 = single steps only) and VL_BATCH_TB_ROWS override.
  b->pass_k = b->precision == VL_F32 ? 5 : 3;
  if (const char* e = std::getenv("VL_BATCH_PASS_STEPS"))
    b->pass_k = std::min(std::max(std::atoi(e), 1), b->precision == VL_F32 ? TB_MAX_K : 4);
  b->pass_rows = b->precision == VL_F16 ? 24 : (b->precision == VL_F32 ? 14 : 16);
  if (const char* e = std::getenv("VL_BATCH_TB_ROWS")) b->pass_rows = std::max(1, std::atoi(e));
  // 8 waves per workgroup (FP64: 4); on the 512-case sweep FP32 in 4-wave workgroups, which
  // find more lockstep groups among narrow cases, ran 8 % slower, and tasks of 32 rows 7 %.
  b->pass_waves = b->precision == VL_F64 ? 4 : 8;
  std::vector<PassPlan> plans;
  if (b->pass_k > 1) {
    const int K = b->pass_k, V = b->precision == VL_F16 ? 2 : 1, halo = tb_halo(K, V);
    // strips on the grid from x = 32, as in the single domain (starting them at the first
    // usable column instead, which drops the narrow first strip of each row band, ran 0.5 %
    // slower on the 512-case sweep)
    int x0_grid = std::max(32, halo);
    x0_grid += (x0_grid - halo) % V;  // the strip's first column is a multiple of V
    plans.assign(nc, PassPlan{});
    const unsigned nth = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    std::vector<std::thread> pool;
    for (unsigned w = 0; w < nth; ++w)
      pool.emplace_back([&, w] {
        for (size_t i = w; i < nc; i += nth) {
          const Tables& tab = b->tabs[i];
          st[i] = plan_passes(tab.flag, (int)tab.nx, (int)tab.ny, b->hcases[i].g.pitch, K, V, b->pass_rows, x0_grid,
                              0, plans[i]);
          if (st[i] != VL_OK) err[i] = vl_last_error();
        }
      });
    for (auto& th : pool) th.join();
    for (size_t i = 0; i < nc; ++i)
      if (st[i] != VL_OK) return set_error(st[i], "case " + std::to_string(i) + ": " + err[i]);
  }
  lap(1);
  // The region kernel is bit-identical but measured slower on the 512-case sweep (22.8 cases/s
  // after the pass, 21.7 beside the next far pass, against 25.4 with the per-step kernels): one
  // workgroup per case cannot hide the partial-line traffic of the nodes beside the inlet and
  // outlet, and beside the far pass that traffic slows the pass more than the overlap hides.
  b->region = false;
  if (const char* e = std::getenv("VL_BATCH_REGION")) b->region = b->pass_k > 1 && e[0] == '1';
  if (const char* e = std::getenv("VL_BATCH_REGION_T")) b->region_threads = std::atoi(e);
  if (b->region_threads != 256 && b->region_threads != 1024) b->region_threads = 512;
  if (const char* e = std::getenv("VL_BATCH_REGION0")) b->region0_side = std::string(e) != "after";
  if (const char* e = std::getenv("VL_BATCH_OVERLAP")) b->overlap = e[0] != '0';
  // Band side arrays: with K-step passes and the per-step kernels.  A case's bands are its K
  // columns beside the inlet and the outlet, which no pass task can cover (a covered node's
  // K - 1 columns on each side are interior, and column 1 and nx - 2 are boundary columns); a
  // case too narrow for two separate bands keeps every node in the pools.
  b->band = b->pass_k > 1 && !b->region;
  if (const char* e = std::getenv("VL_BATCH_BAND")) b->band = b->band && e[0] != '0';
  int64_t side_elems = 0, bidx_len = 0;
  for (size_t i = 0; i < nc; ++i) {
    BatchCase& c = b->hcases[i];
    c.band_w = b->band && c.g.nx >= 2 * b->pass_k + 4 ? b->pass_k : 0;
    c.band_sny = (int32_t)rup(c.g.ny, 64);
    c.side_off = side_elems;
    c.bidx_off = bidx_len;
    if (c.band_w > 0) {
      side_elems += (int64_t)Q * 2 * c.band_w * c.band_sny;
      bidx_len += (int64_t)2 * c.band_w * c.band_sny;
    }
  }
  if (b->band)
    for (int L = 0; L < b->pass_k; ++L) {
      void*& p = L == 0 ? b->bside0 : b->bside[L - 1];
      VL_HIP(hipMalloc(&p, (size_t)std::max<int64_t>(side_elems, 1) * b->elem));
      b->device_bytes += std::max<int64_t>(side_elems, 1) * (int64_t)b->elem;
    }
  b->nbuf = b->pass_k > 1 ? b->pass_k + 1 : 2;
  for (int j = 0; j < b->nbuf; ++j) VL_HIP(hipMalloc(&b->pool[j], (size_t)pop_elems * b->elem));
  b->device_bytes += b->nbuf * pop_elems * (int64_t)b->elem;
  // flags (pitched; padding columns are "not fluid"), unless the set-up kernels wrote them
  if (!b->setup_on_device) {
    std::vector<uint8_t> fl((size_t)flag_bytes, 2);
    for (size_t i = 0; i < nc; ++i) {
      const Tables& tab = b->tabs[i];
      const BatchCase& c = b->hcases[i];
      for (int64_t y = 0; y < tab.ny; ++y)
        std::memcpy(&fl[(size_t)(c.flag_off + y * c.g.pitch)], &tab.flag[(size_t)(y * tab.nx)], (size_t)tab.nx);
    }
    VL_TRYB(dupload(b, b->dflags, fl));
  }
  // packed boundary tables
  if (b->setup_on_device) {
    std::vector<int32_t> gfirst;
    std::vector<int64_t> o_g(nc), o_fc(nc), o_pv(nc);
    int64_t fc_len = 0, pv_len = 0;
    int64_t ob = 0, obf = 0, ol = 0, op = 0;
    std::vector<int64_t> o_b(nc), o_bf(nc), o_l(nc), o_p(nc);
    for (size_t i = 0; i < nc; ++i) {
      const SetupCounts& k = dcount[i];
      o_b[i] = ob;
      o_bf[i] = obf;
      o_l[i] = ol;
      o_p[i] = op;
      ob += k.nb;
      obf += k.nb + 1;
      ol += k.nl;
      op += k.npb;
      o_g[i] = (int64_t)gfirst.size();
      gfirst.push_back(0);  // the body's force group, when it has force links
      if (k.nslots > 0) gfirst.push_back((int32_t)k.nslots);
      o_fc[i] = fc_len;
      fc_len += 4 * std::max<int64_t>(k.nslots, 1);
      o_pv[i] = pv_len;
      pv_len += 2 * std::max<int64_t>(k.npb, 1);
    }
    int32_t* d_gf;
    double *d_fc, *d_pv, *d_drho;
    VL_TRYB(dupload(b, d_gf, gfirst));
    std::vector<double> zeros_fc((size_t)fc_len, 0.0), zeros_pv((size_t)pv_len, 0.0), zeros_d(nc, 0.0);
    VL_TRYB(dupload(b, d_fc, zeros_fc));
    VL_TRYB(dupload(b, d_pv, zeros_pv));
    VL_TRYB(dupload(b, d_drho, zeros_d));
    for (size_t i = 0; i < nc; ++i) {
      const SetupCounts& k = dcount[i];
      DevTables& t = b->hcases[i].t;
      t.nb = (int32_t)k.nb;
      t.nl = (int32_t)k.nl;
      t.np = (int32_t)k.npb;
      t.npb = (int32_t)k.npb;
      t.nslots = (int32_t)k.nslots;
      t.ngroups = k.nslots > 0 ? 1 : 0;
      t.bx = dtab.bx + o_b[i];
      t.by = dtab.by + o_b[i];
      t.bfirst = dtab.bfirst + o_bf[i];
      t.dir = dtab.dir + o_l[i];
      t.kind = dtab.kind + o_l[i];
      t.vgroup = dtab.vgroup + o_l[i];
      t.lflags = dtab.lflags + o_l[i];
      t.lnx = dtab.lnx + o_l[i];
      t.lny = dtab.lny + o_l[i];
      t.fslot = dtab.fslot + o_l[i];
      t.q = dtab.q + o_l[i];
      t.uwx = dtab.uwx + o_l[i];
      t.uwy = dtab.uwy + o_l[i];
      t.cu6 = dtab.cu6 + o_l[i];
      t.rho = dtab.rho + o_l[i];
      t.px = dtab.px + o_p[i];
      t.py = dtab.py + o_p[i];
      t.pnx = dtab.pnx + o_p[i];
      t.pny = dtab.pny + o_p[i];
      t.gfirst = d_gf + o_g[i];
      t.fcontrib = d_fc + o_fc[i];
      t.pvel = d_pv + o_pv[i];
      t.drho = d_drho + i;
    }
  } else {
    std::vector<int32_t> bx, by, bfirst, fslot, px, py, gfirst;
    std::vector<uint8_t> dir, kind, vgroup, lflags;
    std::vector<int8_t> lnx, lny, pnx, pny;
    std::vector<double> q, uwx, uwy, cu6, rho;
    std::vector<int64_t> o_b(nc), o_bf(nc), o_l(nc), o_p(nc), o_g(nc), o_fc(nc), o_pv(nc);
    int64_t fc_len = 0, pv_len = 0;
    for (size_t i = 0; i < nc; ++i) {
      const Tables& tab = b->tabs[i];
      o_b[i] = (int64_t)bx.size();
      o_bf[i] = (int64_t)bfirst.size();
      o_l[i] = (int64_t)dir.size();
      o_p[i] = (int64_t)px.size();
      o_g[i] = (int64_t)gfirst.size();
      bx.insert(bx.end(), tab.bnode_x.begin(), tab.bnode_x.end());
      by.insert(by.end(), tab.bnode_y.begin(), tab.bnode_y.end());
      bfirst.insert(bfirst.end(), tab.bnode_first.begin(), tab.bnode_first.end());  // case-local link ids
      dir.insert(dir.end(), tab.l_dir.begin(), tab.l_dir.end());
      kind.insert(kind.end(), tab.l_kind.begin(), tab.l_kind.end());
      vgroup.insert(vgroup.end(), tab.l_vgroup.begin(), tab.l_vgroup.end());
      lflags.insert(lflags.end(), tab.l_flags.begin(), tab.l_flags.end());
      lnx.insert(lnx.end(), tab.l_nx.begin(), tab.l_nx.end());
      lny.insert(lny.end(), tab.l_ny.begin(), tab.l_ny.end());
      fslot.insert(fslot.end(), tab.l_fslot.begin(), tab.l_fslot.end());
      q.insert(q.end(), tab.l_q.begin(), tab.l_q.end());
      uwx.insert(uwx.end(), tab.l_uwx.begin(), tab.l_uwx.end());
      uwy.insert(uwy.end(), tab.l_uwy.begin(), tab.l_uwy.end());
      cu6.insert(cu6.end(), tab.l_cu6.begin(), tab.l_cu6.end());
      rho.insert(rho.end(), tab.l_rho.begin(), tab.l_rho.end());
      px.insert(px.end(), tab.p_x.begin(), tab.p_x.end());
      py.insert(py.end(), tab.p_y.begin(), tab.p_y.end());
      pnx.insert(pnx.end(), tab.p_nx.begin(), tab.p_nx.end());
      pny.insert(pny.end(), tab.p_ny.begin(), tab.p_ny.end());
      gfirst.insert(gfirst.end(), tab.f_group_first.begin(), tab.f_group_first.end());
      o_fc[i] = fc_len;
      fc_len += 4 * std::max<int64_t>((int64_t)tab.f_link.size(), 1);
      o_pv[i] = pv_len;
      pv_len += 2 * std::max<int64_t>((int64_t)tab.p_x.size(), 1);
    }
    int32_t *d_bx, *d_by, *d_bf, *d_fs, *d_px, *d_py, *d_gf;
    uint8_t *d_dir, *d_kind, *d_vg, *d_lf;
    int8_t *d_lnx, *d_lny, *d_pnx, *d_pny;
    double *d_q, *d_uwx, *d_uwy, *d_cu6, *d_rho, *d_fc, *d_pv, *d_drho;
    VL_TRYB(dupload(b, d_bx, bx));
    VL_TRYB(dupload(b, d_by, by));
    VL_TRYB(dupload(b, d_bf, bfirst));
    VL_TRYB(dupload(b, d_fs, fslot));
    VL_TRYB(dupload(b, d_px, px));
    VL_TRYB(dupload(b, d_py, py));
    VL_TRYB(dupload(b, d_gf, gfirst));
    VL_TRYB(dupload(b, d_dir, dir));
    VL_TRYB(dupload(b, d_kind, kind));
    VL_TRYB(dupload(b, d_vg, vgroup));
    VL_TRYB(dupload(b, d_lf, lflags));
    VL_TRYB(dupload(b, d_lnx, lnx));
    VL_TRYB(dupload(b, d_lny, lny));
    VL_TRYB(dupload(b, d_pnx, pnx));
    VL_TRYB(dupload(b, d_pny, pny));
    VL_TRYB(dupload(b, d_q, q));
    VL_TRYB(dupload(b, d_uwx, uwx));
    VL_TRYB(dupload(b, d_uwy, uwy));
    VL_TRYB(dupload(b, d_cu6, cu6));
    VL_TRYB(dupload(b, d_rho, rho));
    std::vector<double> zeros_fc((size_t)fc_len, 0.0), zeros_pv((size_t)pv_len, 0.0), zeros_d(nc, 0.0);
    VL_TRYB(dupload(b, d_fc, zeros_fc));
    VL_TRYB(dupload(b, d_pv, zeros_pv));
    VL_TRYB(dupload(b, d_drho, zeros_d));
    for (size_t i = 0; i < nc; ++i) {
      const Tables& tab = b->tabs[i];
      DevTables& t = b->hcases[i].t;
      t.nb = (int32_t)tab.bnode_x.size();
      t.nl = (int32_t)tab.l_dir.size();
      t.np = (int32_t)tab.p_x.size();
      t.npb = npb[i];
      t.nslots = (int32_t)tab.f_link.size();
      t.ngroups = tab.force_groups;
      if (t.ngroups > 1) return set_error(VL_EINVAL, "batched cases have one force group (the body)");
      t.bx = d_bx + o_b[i];
      t.by = d_by + o_b[i];
      t.bfirst = d_bf + o_bf[i];
      t.dir = d_dir + o_l[i];
      t.kind = d_kind + o_l[i];
      t.vgroup = d_vg + o_l[i];
      t.lflags = d_lf + o_l[i];
      t.lnx = d_lnx + o_l[i];
      t.lny = d_lny + o_l[i];
      t.fslot = d_fs + o_l[i];
      t.q = d_q + o_l[i];
      t.uwx = d_uwx + o_l[i];
      t.uwy = d_uwy + o_l[i];
      t.cu6 = d_cu6 + o_l[i];
      t.rho = d_rho + o_l[i];
      t.px = d_px + o_p[i];
      t.py = d_py + o_p[i];
      t.pnx = d_pnx + o_p[i];
      t.pny = d_pny + o_p[i];
      t.gfirst = d_gf + o_g[i];
      t.fcontrib = d_fc + o_fc[i];
      t.pvel = d_pv + o_pv[i];
      t.drho = d_drho + i;
    }
  }
  // ---- tile tables: cases in descending step order ----
  std::vector<int> order(nc);
  for (size_t i = 0; i < nc; ++i) order[i] = (int)i;
  std::stable_sort(order.begin(), order.end(),
                   [&](int a, int c) { return b->hcases[a].steps > b->hcases[c].steps; });
  std::vector<int32_t> bc, bb0, bact((size_t)b->max_steps + 2, 0);
  std::vector<int32_t> bnd_end(nc, 0);
  for (int ci : order) {
    BatchCase& c = b->hcases[ci];
    c.bnd_tile0 = (int32_t)bc.size();
    const int32_t nb = c.t.nb;
    for (int32_t b0 = 0; b0 < std::max(nb, 1); b0 += BT) {  // at least one tile: it owns the force history
      bc.push_back(ci);
      bb0.push_back(b0);
    }
    bnd_end[ci] = (int32_t)bc.size();
  }
  b->nbnd_tiles = (int)bc.size();
  // boundary tiles run for t <= steps (step `steps` only reduces the last step's forces)
  for (int64_t tt = 0; tt <= b->max_steps + 1; ++tt) {
    int32_t n = 0;
    for (int ci : order)
      if (b->hcases[ci].steps >= tt) n = bnd_end[ci];
    bact[(size_t)tt] = n;
  }
  VL_TRYB(dupload(b, b->d_bcase, bc));
  VL_TRYB(dupload(b, b->d_bb0, bb0));
  VL_TRYB(dupload(b, b->d_bact, bact));
  b->h_bact = bact;
  // slot-major link copy: LINK_SLOTS slots of every boundary node of every case
  if (const char* e = std::getenv("VL_BATCH_LINK_SLOTS")) b->link_slots = e[0] != '0';
  if (b->link_slots) {
    int64_t nbt = 0;
    for (size_t i = 0; i < nc; ++i) nbt += b->hcases[i].t.nb;
    const int64_t n = std::max<int64_t>(nbt, 1) * LINK_SLOTS;
    uint32_t* m;
    int32_t* fs;
    double *q, *cu6, *rho, *uwx, *uwy;
    VL_TRYB(dalloc(b, m, (size_t)n));
    VL_TRYB(dalloc(b, fs, (size_t)n));
    VL_TRYB(dalloc(b, q, (size_t)n));
    VL_TRYB(dalloc(b, cu6, (size_t)n));
    VL_TRYB(dalloc(b, rho, (size_t)n));
    VL_TRYB(dalloc(b, uwx, (size_t)n));
    VL_TRYB(dalloc(b, uwy, (size_t)n));
    void* arrs[7] = {m, fs, q, cu6, rho, uwx, uwy};
    for (int k = 0; k < 7; ++k) b->lslot[k] = arrs[k];
    int64_t off = 0;
    for (size_t i = 0; i < nc; ++i) {
      LinkSlots& ls = b->hcases[i].ls;
      ls.meta = m + off;
      ls.fslot = fs + off;
      ls.q = q + off;
      ls.cu6 = cu6 + off;
      ls.rho = rho + off;
      ls.uwx = uwx + off;
      ls.uwy = uwy + off;
      ls.stride = std::max<int64_t>(nbt, 1);
      off += b->hcases[i].t.nb;
    }
  }
  VL_TRYB(dupload(b, b->dcases, b->hcases));
  VL_TRYB(dalloc(b, b->d_status, nc));
  VL_TRYB(dalloc(b, b->d_done, nc));
  VL_HIP(hipMemset(b->d_status, 0, nc * sizeof(int32_t)));
  VL_HIP(hipMemset(b->d_done, 0, nc * sizeof(int64_t)));
  VL_TRYB(dalloc(b, b->d_tbase, 2));
  VL_HIP(hipMemset(b->d_tbase, 0, 2 * sizeof(int64_t)));
  VL_TRYB(dalloc(b, b->d_hist, (size_t)std::max<int64_t>(hist_len, 2)));
  VL_HIP(hipMemset(b->d_hist, 0, (size_t)std::max<int64_t>(hist_len, 2) * sizeof(double)));
  if (kick_tab.empty()) kick_tab.push_back(0.0);
  VL_TRYB(dupload(b, b->d_kick, kick_tab));
  // ---- initial state, in every pool (two at a time), all cases per launch (rows of 64
  // workgroups, up to 65535 cases a launch) ----
  for (size_t c0 = 0; c0 < nc; c0 += 65535) {
    const unsigned rows = (unsigned)std::min<size_t>(65535, nc - c0);
    for (int j = 0; j < b->nbuf; j += 2) {
      void* pa = b->pool[j];
      void* pb = b->pool[std::min(j + 1, b->nbuf - 1)];
      switch (b->precision) {
        case VL_F64:
          hipLaunchKernelGGL(k_batch_init<StoreF64>, dim3(64, rows), dim3(BT), 0, b->stream, (double*)pa, (double*)pb,
                             b->dflags, b->dcases, (int)c0);
          break;
        case VL_F16:
          hipLaunchKernelGGL(k_batch_init<StoreF16>, dim3(64, rows), dim3(BT), 0, b->stream, (_Float16*)pa,
                             (_Float16*)pb, b->dflags, b->dcases, (int)c0);
          break;
        default:
          hipLaunchKernelGGL(k_batch_init<StoreF32>, dim3(64, rows), dim3(BT), 0, b->stream, (float*)pa, (float*)pb,
                             b->dflags, b->dcases, (int)c0);
      }
      VL_HIP(hipGetLastError());
    }
    // pressure-node velocities of the initial state: the boundary update writes them from then on
    switch (b->precision) {
      case VL_F64:
        hipLaunchKernelGGL(k_batch_outlet_prime<StoreF64>, dim3(4, rows), dim3(BT), 0, b->stream,
                           (const double*)b->pool[0], b->dcases, (int)c0);
        break;
      case VL_F16:
        hipLaunchKernelGGL(k_batch_outlet_prime<StoreF16>, dim3(4, rows), dim3(BT), 0, b->stream,
                           (const _Float16*)b->pool[0], b->dcases, (int)c0);
        break;
      default:
        hipLaunchKernelGGL(k_batch_outlet_prime<StoreF32>, dim3(4, rows), dim3(BT), 0, b->stream,
                           (const float*)b->pool[0], b->dcases, (int)c0);
    }
    VL_HIP(hipGetLastError());
  }
  if (b->link_slots)
    for (size_t c0 = 0; c0 < nc; c0 += 65535) {
      const unsigned rows = (unsigned)std::min<size_t>(65535, nc - c0);
      hipLaunchKernelGGL(k_link_slots, dim3(16, rows), dim3(BT), 0, b->stream, b->dcases, (int)c0,
                         (uint32_t*)b->lslot[0], (int32_t*)b->lslot[1], (double*)b->lslot[2], (double*)b->lslot[3],
                         (double*)b->lslot[4], (double*)b->lslot[5], (double*)b->lslot[6]);
      VL_HIP(hipGetLastError());
    }
  if (b->band) {
    VL_TRYB(dalloc(b, b->d_bidx, (size_t)std::max<int64_t>(bidx_len, 1)));
    for (size_t c0 = 0; c0 < nc; c0 += 65535) {
      const unsigned rows = (unsigned)std::min<size_t>(65535, nc - c0);
      hipLaunchKernelGGL(k_band_map_init, dim3(16, rows), dim3(BT), 0, b->stream, b->d_bidx, b->dflags, b->dcases,
                         (int)c0);
      hipLaunchKernelGGL(k_band_map_bnd, dim3(16, rows), dim3(BT), 0, b->stream, b->d_bidx, b->dcases, (int)c0);
      VL_HIP(hipGetLastError());
    }
  }
  VL_HIP(hipStreamSynchronize(b->stream));
  lap(2);
  // ---- autotune the bulk tile shape and the grid policy on the real batch ----
  if (const char* e = std::getenv("VL_BATCH_SERIAL")) b->serial = e[0] == '1';
  b->step_overlap = false;
  if (const char* e = std::getenv("VL_BATCH_STEP_OVERLAP")) b->step_overlap = b->band && b->third && !b->serial && e[0] == '1';
  hipDeviceProp_t prop;
  VL_HIP(hipGetDeviceProperties(&prop, 0));
  const int cus = std::max(1, prop.multiProcessorCount);
  b->grid_bnd = std::max(1, b->nbnd_tiles);  // one workgroup per boundary tile
  if (const char* e = std::getenv("VL_BATCH_GRID_BND")) b->grid_bnd = std::max(1, std::min(b->nbnd_tiles, std::atoi(e)));
  const int shapes[3][2] = {{6, 4}, {7, 2}, {8, 1}};  // (log2 TW, TH)
  const int mults[4] = {0, 4, 8, 16};  // 0: one workgroup per tile; else workgroups per CU
  double best = 1e30;
  std::vector<int32_t> kc, ky, kx, kend;
  hipEvent_t e0, e1;
  VL_HIP(hipEventCreate(&e0));
  VL_HIP(hipEventCreate(&e1));
  int best_s = 0, best_m = 1;
  // The persistent grids lost to one workgroup per tile on every batch measured (the 512-case
  // sweep: 1.78-1.82 ms against 2.0-2.6 ms a trial step), so by default only the three tile
  // shapes are timed, one workgroup per tile, over two steps each; VL_BATCH_AUTOTUNE=full times all
  // twelve policies over four.
  // With K-step passes the bulk tiles only carry the cases that end inside a group, a small
  // share of the work, and the shape made no measurable difference, so the trials are skipped
  // (128 x 2 tiles) unless VL_BATCH_AUTOTUNE asks for them ("on", or "full").
  bool full = false, tune = b->pass_k == 1;
  if (const char* e = std::getenv("VL_BATCH_AUTOTUNE")) {
    full = std::string(e) == "full";
    tune = full || std::string(e) == "on";
  }
  if (!tune) best_s = 1;
  const int nmult = full ? 4 : 1, reps = full ? 4 : 2;
  for (int si = 0; tune && si < 3; ++si) {
    bulk_tiles(b, order, 1 << shapes[si][0], shapes[si][1], kc, ky, kx, kend);
    int32_t *tkc, *tky, *tkx, *tact;
    std::vector<int32_t> act1(1, (int32_t)kc.size());
    VL_TRYB(dupload(b, tkc, kc));
    VL_TRYB(dupload(b, tky, ky));
    VL_TRYB(dupload(b, tkx, kx));
    VL_TRYB(dupload(b, tact, act1));
    for (int mi = 0; mi < nmult; ++mi) {
      const int grid = mults[mi] == 0 ? std::max(1, (int)kc.size())
                                      : std::max(1, std::min((int)kc.size(), cus * mults[mi]));
      // whole trial steps 0 (boundary beside bulk), A -> B from the same A every time
      VL_TRYB(trial_step(b, grid, shapes[si][0], tkc, tky, tkx, tact));
      VL_HIP(hipEventRecord(e0, b->stream));
      for (int r = 0; r < reps; ++r) VL_TRYB(trial_step(b, grid, shapes[si][0], tkc, tky, tkx, tact));
      VL_HIP(hipEventRecord(e1, b->stream));
      VL_HIP(hipEventSynchronize(e1));
      float ms = 0.0f;
      VL_HIP(hipEventElapsedTime(&ms, e0, e1));
      b->autotune_ms[si][mi] = ms / reps;
      if (ms < best) {
        best = ms;
        best_s = si;
        best_m = mi;
      }
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (const char* e = std::getenv("VL_BATCH_TILE")) {  // force a shape: 0, 1, 2 = 64x4, 128x2, 256x1
    const int si = std::atoi(e);
    if (si >= 0 && si < 3) best_s = si;
  }
  b->tw_log2 = shapes[best_s][0];
  b->th = shapes[best_s][1];
  bulk_tiles(b, order, 1 << b->tw_log2, b->th, kc, ky, kx, kend);
  b->nbulk_tiles = (int)kc.size();
  b->bulk_mult = mults[best_m];
  if (const char* e = std::getenv("VL_BATCH_BULK_MULT")) b->bulk_mult = std::max(0, std::atoi(e));
  b->grid_bulk = b->bulk_mult == 0 ? std::max(1, b->nbulk_tiles)
                                   : std::max(1, std::min(b->nbulk_tiles, cus * b->bulk_mult));
  make_ladder(b);
  std::vector<int32_t> kact((size_t)b->max_steps + 2, 0);
  for (int64_t tt = 0; tt <= b->max_steps + 1; ++tt) {
    int32_t n = 0;
    for (int ci : order)
      if (b->hcases[ci].steps > tt) n = kend[ci];
    kact[(size_t)tt] = n;
  }
  VL_TRYB(dupload(b, b->d_kcase, kc));
  VL_TRYB(dupload(b, b->d_ky0, ky));
  VL_TRYB(dupload(b, b->d_kx0, kx));
  VL_TRYB(dupload(b, b->d_kact, kact));
  b->h_kact = std::move(kact);
  // ---- pass tasks and listed tiles, cases in descending step order ----
  {
    // three task lists: all tasks (one launch per group), and the far and near tasks (two
    // launches, when the region kernel overlaps the next group's far pass); in each, a case's
    // tasks in (first row, first column) order, so that a workgroup's waves sweep adjacent
    // strips of the same rows
    std::vector<int4> task, ftask, ntask;
    std::vector<int32_t> tcase, ftcase, ntcase, lcase, lfirst, lcount;
    std::vector<uint32_t> lnodes;
    std::vector<int32_t> tend(nc, 0), ftend(nc, 0), ntend(nc, 0), lend(nc, 0), clfirst(nc, 0), clcount(nc, 0),
        bend(nc, 0), cend(nc, 0);
    std::vector<int4> btile;
    std::vector<int2> ctile;
    auto by_row = [](const int4& a, const int4& c) { return a.y != c.y ? a.y < c.y : a.x < c.x; };
    if (b->pass_k > 1)
      for (int ci : order) {
        PassPlan& pl = plans[(size_t)ci];
        std::vector<int4> far(pl.tasks.begin(), pl.tasks.begin() + pl.nfar), near(pl.tasks.begin() + pl.nfar, pl.tasks.end());
        std::sort(pl.tasks.begin(), pl.tasks.end(), by_row);
        std::sort(far.begin(), far.end(), by_row);
        std::sort(near.begin(), near.end(), by_row);
        for (const int4& tk : pl.tasks) {
          task.push_back(tk);
          tcase.push_back(ci);
        }
        for (const int4& tk : far) {
          ftask.push_back(tk);
          ftcase.push_back(ci);
        }
        for (const int4& tk : near) {
          ntask.push_back(tk);
          ntcase.push_back(ci);
        }
        ftend[(size_t)ci] = (int32_t)ftask.size();
        ntend[(size_t)ci] = (int32_t)ntask.size();
        // with side arrays the band kernel takes the band columns' nodes: the listed tiles hold
        // the others
        const BatchCase& bc_ = b->hcases[(size_t)ci];
        std::vector<uint32_t> lst;
        lst.reserve(pl.listed.size());
        for (uint32_t v : pl.listed) {
          const int x = (int)(v & 0xffffu);
          if (bc_.band_w > 0 && (x <= bc_.band_w || x >= bc_.g.nx - 1 - bc_.band_w)) continue;
          lst.push_back(v);
        }
        for (size_t k = 0; k < lst.size(); k += (size_t)LISTED_NPT * BT) {
          lcase.push_back(ci);
          lfirst.push_back((int32_t)(lnodes.size() + k));
          lcount.push_back((int32_t)std::min<size_t>((size_t)LISTED_NPT * BT, lst.size() - k));
        }
        clfirst[(size_t)ci] = (int32_t)lnodes.size();
        clcount[(size_t)ci] = (int32_t)lst.size();
        lnodes.insert(lnodes.end(), lst.begin(), lst.end());
        if (bc_.band_w > 0) {
          for (int c = 0; c < 2 * bc_.band_w; ++c)
            for (int y0 = 0; y0 < bc_.g.ny; y0 += BT) btile.push_back(int4{ci, c, y0, 0});
          for (int y0 = 0; y0 < bc_.g.ny; y0 += COPY_ROWS) ctile.push_back(int2{ci, y0});
        }
        bend[(size_t)ci] = (int32_t)btile.size();
        cend[(size_t)ci] = (int32_t)ctile.size();
        tend[(size_t)ci] = (int32_t)task.size();
        lend[(size_t)ci] = (int32_t)lcase.size();
        b->pass_nodes += pl.covered;
        b->listed_nodes += (int64_t)pl.listed.size();
      }
    // per group start t: the tasks, listed tiles and bulk tiles of the cases with steps >= t + K
    const size_t nt = (size_t)b->max_steps + 2;
    vl_batch::TaskList& tl = b->tasks;
    tl.h_n_at.assign(nt, 0);
    b->tasks_far.h_n_at.assign(nt, 0);
    b->tasks_near.h_n_at.assign(nt, 0);
    b->h_ln_at.assign(nt, 0);
    b->h_kfirst_at.assign(nt, 0);
    b->h_rn_at.assign(nt, 0);
    b->h_bfirst_at.assign(nt, 0);
    b->h_nband_at.assign(nt, 0);
    b->h_ncopy_at.assign(nt, 0);
    if (b->pass_k > 1)
      for (size_t tt = 0; tt < nt; ++tt) {
        int32_t ncase = 0;
        for (int ci : order) {
          if (b->hcases[(size_t)ci].steps < (int64_t)tt + b->pass_k) break;
          tl.h_n_at[tt] = tend[(size_t)ci];
          b->tasks_far.h_n_at[tt] = ftend[(size_t)ci];
          b->tasks_near.h_n_at[tt] = ntend[(size_t)ci];
          b->h_ln_at[tt] = lend[(size_t)ci];
          b->h_kfirst_at[tt] = kend[(size_t)ci];
          b->h_bfirst_at[tt] = bnd_end[(size_t)ci];
          b->h_nband_at[tt] = bend[(size_t)ci];
          b->h_ncopy_at[tt] = cend[(size_t)ci];
          b->h_rn_at[tt] = ++ncase;
        }
      }
    // Lockstep: the waves of a workgroup sweep in step when its tasks belong to one case and
    // cover the same rows (so they meet the same barriers, and a retired case stops them all).
    // A group's pass launches the prefix of h_n_at[t] tasks of a list from its task 0, in
    // workgroups of W.
    auto finish_list = [&](vl_batch::TaskList& l, std::vector<int4>& tk, std::vector<int32_t>& tc) -> vl_status {
      TaskRanges ranges;
      std::vector<int32_t> ends(l.h_n_at.begin(), l.h_n_at.end());
      ends.push_back((int32_t)tk.size());
      std::sort(ends.begin(), ends.end());
      ends.erase(std::unique(ends.begin(), ends.end()), ends.end());
      for (int32_t e : ends) ranges.emplace_back(0, e);
      mark_lockstep(tk, tc.data(), ranges, b->pass_waves);
      VL_TRYB(check_lockstep(tk, tc.data(), ranges, b->pass_waves));
      l.n = (int)tk.size();
      VL_TRYB(dupload(b, l.task, tk));
      VL_TRYB(dupload(b, l.tcase, tc));
      VL_TRYB(dupload(b, l.n_at, l.h_n_at));
      return VL_OK;
    };
    VL_TRYB(finish_list(tl, task, tcase));
    VL_TRYB(finish_list(b->tasks_far, ftask, ftcase));
    VL_TRYB(finish_list(b->tasks_near, ntask, ntcase));
    b->nlisted_tiles = (int)lcase.size();
    VL_TRYB(dupload(b, b->d_lcase, lcase));
    VL_TRYB(dupload(b, b->d_lfirst, lfirst));
    VL_TRYB(dupload(b, b->d_lcount, lcount));
    VL_TRYB(dupload(b, b->d_lnodes, lnodes));
    VL_TRYB(dupload(b, b->d_ln_at, b->h_ln_at));
    VL_TRYB(dupload(b, b->d_kfirst_at, b->h_kfirst_at));
    std::vector<int32_t> rcase(order.begin(), order.end());
    VL_TRYB(dupload(b, b->d_rcase, rcase));
    VL_TRYB(dupload(b, b->d_rn_at, b->h_rn_at));
    VL_TRYB(dupload(b, b->d_bfirst_at, b->h_bfirst_at));
    VL_TRYB(dupload(b, b->d_clfirst, clfirst));
    VL_TRYB(dupload(b, b->d_clcount, clcount));
    b->nband_tiles = (int)btile.size();
    VL_TRYB(dupload(b, b->d_btile, btile));
    VL_TRYB(dupload(b, b->d_nband_at, b->h_nband_at));
    b->ncopy_tiles = (int)ctile.size();
    VL_TRYB(dupload(b, b->d_ctile, ctile));
    VL_TRYB(dupload(b, b->d_ncopy_at, b->h_ncopy_at));
  }
  b->side0_ok = false;
  b->use_graphs = true;
  if (const char* e = std::getenv("VL_GRAPHS")) b->use_graphs = e[0] != '0';
  lap(3);
  b->built = true;
  return VL_OK;
}

// Steps per replayed window: GRAPH_W single steps, or the whole groups of K steps that fit.
int window_steps(const vl_batch* b) { return b->pass_k > 1 ? GRAPH_W / b->pass_k * b->pass_k : GRAPH_W; }

vl_status window_graph(vl_batch* b, hipGraphExec_t* out) {
  const int wsteps = window_steps(b);
  const int K = b->pass_k;
  GroupGrids gg{0, 0, 0, 0, 0, 0, 0};
  const int grid = K > 1 ? (gg = group_grids_at(b, b->t, wsteps / K)).bulk : bulk_grid_at(b, b->t);
  const bool ok0 = b->side0_ok;
  const auto key =
      std::make_tuple(b->cur, grid, gg.pass, gg.listed, gg.bnd, gg.pass_near, gg.band, gg.copy, (int)ok0);
  auto found = b->graphs.find(key);
  if (found != b->graphs.end()) {
    *out = found->second;
    return VL_OK;
  }
  const int cur0 = b->cur;
  const int64_t t0 = b->t;
  hipGraph_t g = nullptr;
  VL_HIP(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
  vl_status st = VL_OK;
  if (K > 1) {
    for (int k = 0; k < wsteps && st == VL_OK; k += K) st = batch_group(b, b->d_tbase, k, gg);
    if (st == VL_OK) st = join_side(b);
  } else
    for (int k = 0; k < wsteps && st == VL_OK; ++k) st = batch_step(b, b->d_tbase, k, grid);
  if (st == VL_OK) hipLaunchKernelGGL(k_batch_advance, dim3(1), dim3(1), 0, b->stream, b->d_tbase, (int64_t)wsteps);
  hipError_t e = hipStreamEndCapture(b->stream, &g);
  b->cur = cur0;
  b->t = t0;
  b->ov_open = false;
  b->side0_ok = ok0;
  if (st != VL_OK) {
    if (g) (void)hipGraphDestroy(g);
    return st;
  }
  if (e != hipSuccess) return set_error(VL_EDEVICE, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
  hipGraphExec_t ge = nullptr;
  e = hipGraphInstantiate(&ge, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  if (e != hipSuccess) return set_error(VL_EDEVICE, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
  b->graphs[key] = ge;
  *out = ge;
  return VL_OK;
}

}  // namespace

// ==========================================================================================
// C ABI
// ==========================================================================================
extern "C" {

vl_status vl_batch_create(int precision, vl_batch** out) {
  if (!out) return set_error(VL_EINVAL, "null argument");
  if (precision != VL_F64 && precision != VL_F32 && precision != VL_F16)
    return set_error(VL_EINVAL, "precision must be VL_F64, VL_F32 or VL_F16");
  auto* b = new vl_batch;
  b->precision = precision;
  b->elem = precision == VL_F64 ? 8 : (precision == VL_F32 ? 4 : 2);
  *out = b;
  return VL_OK;
}

int vl_batch_add(vl_batch* b, const vl_channel_spec* spec, vl_shape* body, int64_t steps, double d) {
  if (!b || !spec || !body || b->built || steps < 1 || !(d > 0.0)) {
    if (body) vl_shape_free(body);
    set_error(VL_EINVAL, b && b->built ? "batch already built" : "bad case (spec, body, steps >= 1, d > 0)");
    return -1;
  }
  b->pending.push_back({*spec, body, steps, d});
  return (int)b->pending.size() - 1;
}

vl_status vl_batch_build(vl_batch* b) {
  if (!b) return set_error(VL_EINVAL, "null batch");
  if (b->built) return VL_OK;
  vl_status st = build(b);
  for (auto& p : b->pending)
    if (p.body) {
      vl_shape_free(p.body);
      p.body = nullptr;
    }
  return st;
}

vl_status vl_batch_run(vl_batch* b, int64_t max_steps) {
  if (!b) return set_error(VL_EINVAL, "null batch");
  if (!b->built) VL_TRYB(vl_batch_build(b));
  const int64_t end = b->max_steps + 1;  // the extra step reduces the last step's forces
  int64_t todo = end - b->t;
  if (max_steps >= 0) todo = std::min(todo, max_steps);
  if (todo <= 0) return VL_OK;
  const auto t0 = std::chrono::steady_clock::now();
  int64_t k = 0;
  if (b->timing) {
    // direct launches with events around every kernel
    VL_HIP(hipMemcpyAsync(b->d_tbase, &b->t, sizeof(int64_t), hipMemcpyHostToDevice, b->stream));
    VL_HIP(hipStreamSynchronize(b->stream));
    int pending = 0;
    for (int64_t kk = 0; kk < todo; ++kk) {
      VL_TRYB(batch_step(b, b->d_tbase, (int)kk, bulk_grid_at(b, b->t), &b->tev[(size_t)pending]));
      if (++pending == TIMING_W || kk + 1 == todo) {
        VL_TRYB(read_timing(b, pending));
        pending = 0;
      }
    }
    k = todo;
  }
  const int wsteps = window_steps(b);
  if (b->use_graphs && todo - k >= wsteps) {
    // the window counter holds the global step of the window's first step
    VL_HIP(hipMemcpyAsync(b->d_tbase, &b->t, sizeof(int64_t), hipMemcpyHostToDevice, b->stream));
    VL_HIP(hipStreamSynchronize(b->stream));
    for (; k + wsteps <= todo; k += wsteps) {
      hipGraphExec_t ge;
      VL_TRYB(window_graph(b, &ge));
      VL_HIP(hipGraphLaunch(ge, b->stream));
      b->t += wsteps;
      b->cur = (b->cur + wsteps) % b->nbuf;
      b->side0_ok = b->band && b->pass_k > 1;  // a window of whole groups ends with a group
    }
  }
  if (k < todo) {
    // remaining steps launched directly (whole groups of K steps first); their base is the
    // global step in a device scalar
    VL_HIP(hipStreamSynchronize(b->stream));
    VL_HIP(hipMemcpyAsync(b->d_tbase, &b->t, sizeof(int64_t), hipMemcpyHostToDevice, b->stream));
    VL_HIP(hipStreamSynchronize(b->stream));
    int kk = 0;
    if (b->pass_k > 1) {
      for (; k + b->pass_k <= todo; k += b->pass_k, kk += b->pass_k)
        VL_TRYB(batch_group(b, b->d_tbase, kk, group_grids_at(b, b->t, 1)));
      VL_TRYB(join_side(b));
    }
    for (; k < todo; ++k, ++kk) VL_TRYB(batch_step(b, b->d_tbase, kk, bulk_grid_at(b, b->t)));
  }
  VL_HIP(hipStreamSynchronize(b->stream));
  b->last_run_seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  return VL_OK;
}

vl_status vl_batch_set_timing(vl_batch* b, int on) {
  if (!b) return set_error(VL_EINVAL, "null batch");
  if (on && b->tev.empty()) {
    b->tev.resize(TIMING_W);
    for (auto& e : b->tev) {
      VL_HIP(hipEventCreate(&e.bs));
      VL_HIP(hipEventCreate(&e.be));
      VL_HIP(hipEventCreate(&e.ks));
      VL_HIP(hipEventCreate(&e.ke));
      VL_HIP(hipEventCreate(&e.ej));
    }
  }
  b->timing = on != 0;
  b->tm_bnd = b->tm_bulk = b->tm_step = 0.0;
  b->tm_steps = 0;
  return VL_OK;
}

vl_status vl_batch_get_timing(vl_batch* b, double* boundary_ms, double* bulk_ms, double* step_ms, int64_t* steps) {
  if (!b) return set_error(VL_EINVAL, "null batch");
  if (boundary_ms) *boundary_ms = b->tm_bnd;
  if (bulk_ms) *bulk_ms = b->tm_bulk;
  if (step_ms) *step_ms = b->tm_step;
  if (steps) *steps = b->tm_steps;
  return VL_OK;
}

vl_status vl_batch_info_get(vl_batch* b, vl_batch_info* o) {
  if (!b || !o) return set_error(VL_EINVAL, "null argument");
  std::memset(o, 0, sizeof(*o));
  o->ncases = (int64_t)(b->built ? b->hcases.size() : b->pending.size());
  o->total_nodes = b->total_nodes;
  o->fluid_nodes = b->total_fluid;
  // node updates performed so far: each case's fluid nodes times the steps it has completed (a
  // retired case counts the steps it ran before its drag turned non-finite)
  o->lattice_updates = 0;
  if (b->built && !b->hcases.empty()) {
    VL_HIP(hipStreamSynchronize(b->stream));
    std::vector<int64_t> done(b->hcases.size());
    VL_HIP(hipMemcpy(done.data(), b->d_done, done.size() * sizeof(int64_t), hipMemcpyDeviceToHost));
    for (size_t c = 0; c < done.size(); ++c)
      o->lattice_updates += b->tabs[c].fluid_nodes * std::min<int64_t>(done[c], b->hcases[c].steps);
  }
  o->max_steps = b->max_steps;
  o->t = b->t;
  o->boundary_tiles = b->nbnd_tiles;
  o->bulk_tiles = b->nbulk_tiles;
  o->tile_w = 1 << b->tw_log2;
  o->tile_h = b->th;
  o->grid_bulk = b->grid_bulk;
  o->grid_boundary = b->grid_bnd;
  o->last_run_seconds = b->last_run_seconds;
  o->device_bytes = b->device_bytes;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 4; ++j) o->autotune_ms[i][j] = b->autotune_ms[i][j];
  o->pass_steps = b->pass_k;
  o->pass_rows = b->pass_k > 1 ? b->pass_rows : 0;
  o->pass_tasks = b->tasks.n;
  o->listed_tiles = b->nlisted_tiles;
  o->pass_nodes = b->pass_nodes;
  o->listed_nodes = b->listed_nodes;
  o->setup_on_device = b->setup_on_device ? 1 : 0;
  for (int k = 0; k < 4; ++k) o->setup_ms[k] = b->setup_ms[k];
  return VL_OK;
}

vl_status vl_batch_case_result(vl_batch* b, int id, double* coef, int64_t* steps_done, int* status) {
  if (!b || !b->built || id < 0 || (size_t)id >= b->hcases.size()) return set_error(VL_EINVAL, "bad case id");
  VL_HIP(hipStreamSynchronize(b->stream));
  const BatchCase& c = b->hcases[id];
  int32_t stv = 0;
  int64_t done = 0;
  VL_HIP(hipMemcpy(&stv, b->d_status + id, sizeof(int32_t), hipMemcpyDeviceToHost));
  VL_HIP(hipMemcpy(&done, b->d_done + id, sizeof(int64_t), hipMemcpyDeviceToHost));
  if (steps_done) *steps_done = done;
  if (status) *status = stv;
  if (coef && c.steps > 0)
    VL_HIP(hipMemcpy(coef, b->d_hist + c.hist_off, (size_t)(2 * c.steps) * sizeof(double), hipMemcpyDeviceToHost));
  return VL_OK;
}

vl_status vl_batch_case_info(vl_batch* b, int id, int64_t* nx, int64_t* ny, int64_t* fluid_nodes, int64_t* steps) {
  if (!b || !b->built || id < 0 || (size_t)id >= b->hcases.size()) return set_error(VL_EINVAL, "bad case id");
  const BatchCase& c = b->hcases[id];
  if (nx) *nx = c.g.nx;
  if (ny) *ny = c.g.ny;
  if (fluid_nodes) *fluid_nodes = b->tabs[id].fluid_nodes;
  if (steps) *steps = c.steps;
  return VL_OK;
}

// Populations of one case in the reference layout (FP64 f_i).
vl_status vl_batch_read_case(vl_batch* b, int id, double* f) {
  if (!b || !b->built || id < 0 || (size_t)id >= b->hcases.size() || !f) return set_error(VL_EINVAL, "bad case id");
  VL_HIP(hipStreamSynchronize(b->stream));
  const BatchCase& c = b->hcases[id];
  // A case stops updating after its own step count while the batch goes on, so its latest
  // populations sit in the pool its last update wrote: step k writes pool (k + 1) mod nbuf, and
  // every case starts in pool 0.  (A run ends between groups of K steps, where every node of a
  // case in passes is up to date.)
  const int64_t updates = std::min<int64_t>(b->t, c.steps);
  std::vector<char> host((size_t)(Q * c.g.plane) * b->elem);
  VL_HIP(hipMemcpy(host.data(), (char*)b->pool[updates % b->nbuf] + (size_t)c.pop_off * b->elem, host.size(),
                   hipMemcpyDeviceToHost));
  const int64_t nx = c.g.nx, ny = c.g.ny, n = nx * ny;
  for (int i = 0; i < Q; ++i)
    for (int64_t y = 0; y < ny; ++y)
      for (int64_t x = 0; x < nx; ++x) {
        const int64_t e = i * c.g.plane + y * c.g.pitch + x;
        double v;
        if (b->precision == VL_F64)
          v = ((const double*)host.data())[e];
        else if (b->precision == VL_F32)
          v = wgt(i) + (double)((const float*)host.data())[e];
        else
          v = wgt(i) + (double)((float)((const _Float16*)host.data())[e] * F16_INV);
        f[i * n + y * nx + x] = v;
      }
  return VL_OK;
}

vl_status vl_batch_table_digests(vl_batch* b, int id, uint64_t* out, int n) {
  if (!b || !b->built || id < 0 || (size_t)id >= b->hcases.size() || !out || n < 0)
    return set_error(VL_EINVAL, "bad case id or output");
  VL_HIP(hipStreamSynchronize(b->stream));
  const BatchCase& c = b->hcases[id];
  const DevTables& t = c.t;
  std::vector<uint64_t> d;
  auto digest = [&](const void* dev, size_t bytes) -> vl_status {
    std::vector<unsigned char> h(bytes);
    if (bytes) VL_HIP(hipMemcpy(h.data(), dev, bytes, hipMemcpyDeviceToHost));
    uint64_t x = 0xcbf29ce484222325ull;  // FNV-1a
    for (unsigned char v : h) x = (x ^ v) * 0x100000001b3ull;
    d.push_back(x);
    return VL_OK;
  };
  const size_t nb = (size_t)t.nb, nl = (size_t)t.nl, np = (size_t)t.np;
  VL_TRYB(digest(b->dflags + c.flag_off, (size_t)c.g.pitch * (size_t)c.g.ny));
  VL_TRYB(digest(t.bx, nb * 4));
  VL_TRYB(digest(t.by, nb * 4));
  VL_TRYB(digest(t.bfirst, (nb + 1) * 4));
  VL_TRYB(digest(t.dir, nl));
  VL_TRYB(digest(t.kind, nl));
  VL_TRYB(digest(t.vgroup, nl));
  VL_TRYB(digest(t.lflags, nl));
  VL_TRYB(digest(t.lnx, nl));
  VL_TRYB(digest(t.lny, nl));
  VL_TRYB(digest(t.fslot, nl * 4));
  VL_TRYB(digest(t.q, nl * 8));
  VL_TRYB(digest(t.uwx, nl * 8));
  VL_TRYB(digest(t.uwy, nl * 8));
  VL_TRYB(digest(t.cu6, nl * 8));
  VL_TRYB(digest(t.rho, nl * 8));
  VL_TRYB(digest(t.px, np * 4));
  VL_TRYB(digest(t.py, np * 4));
  VL_TRYB(digest(t.pnx, np));
  VL_TRYB(digest(t.pny, np));
  VL_TRYB(digest(t.gfirst, ((size_t)t.ngroups + 1) * 4));
  const int64_t cnt[7] = {t.nb, t.nl, t.np, t.npb, t.nslots, t.ngroups, b->tabs[id].fluid_nodes};
  uint64_t x = 0xcbf29ce484222325ull;
  for (size_t k = 0; k < sizeof(cnt); ++k) x = (x ^ ((const unsigned char*)cnt)[k]) * 0x100000001b3ull;
  d.push_back(x);
  for (int k = 0; k < std::min<int>(n, (int)d.size()); ++k) out[k] = d[(size_t)k];
  return VL_OK;
}

void vl_batch_destroy(vl_batch* b) { free_batch(b); }

}  // extern "C"
