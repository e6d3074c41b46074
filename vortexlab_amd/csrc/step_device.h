// Device code of one D2Q9 time step, shared by the single-domain solver (sim.hip) and the
// batched solver (batch.hip) so that both run exactly the same arithmetic per node.
#pragma once

#include <hip/hip_runtime.h>

#include "lbm_math.h"
#include "sim_internal.h"

namespace vl {

// ------------------------------------------------------------------------------------------
// Storage classes
// ------------------------------------------------------------------------------------------
// T: stored type; C: compute type; V: nodes per thread of the vectorised bulk (16-byte loads);
// cv / enc convert one stored value to / from the compute type.
struct StoreF64 {
  using T = double;
  using C = double;
  static constexpr bool exact = true;
  static constexpr int V = 2;
  __device__ static inline C cv(T v) { return v; }
  __device__ static inline T enc(C v) { return v; }
  __device__ static inline C ld(const T* p) { return *p; }
  __device__ static inline void st(T* p, C v) { *p = v; }
  __device__ static inline void st_stream(T* p, C v) { __builtin_nontemporal_store(v, p); }
  __device__ static inline C round(C v) { return v; }
};
struct StoreF32 {
  using T = float;
  using C = float;
  static constexpr bool exact = false;
  static constexpr int V = 4;
  __device__ static inline C cv(T v) { return v; }
  __device__ static inline T enc(C v) { return v; }
  __device__ static inline C ld(const T* p) { return *p; }
  __device__ static inline void st(T* p, C v) { *p = v; }
  __device__ static inline void st_stream(T* p, C v) { __builtin_nontemporal_store(v, p); }
  __device__ static inline C round(C v) { return v; }
};
// FP16 deviations scaled by 2^10: keeps |d| ~ 1e-6..1e-1 inside the normal half range.
constexpr float F16_SCALE = 1024.0f;
constexpr float F16_INV = 1.0f / 1024.0f;
// An FP32 value rounded to FP32 and then to FP16 (two roundings, as the source says).  The
// empty asm keeps the compiler from folding the FMA that produced `v` into a mixed-precision
// FMA with an FP16 result (one rounding): that fold happens in some unrollings and not in
// others, and would make the same node's bits depend on the kernel variant.
__device__ inline _Float16 to_f16(float v) {
  asm("" : "+v"(v));
  return (_Float16)(v * F16_SCALE);
}
struct StoreF16 {
  using T = _Float16;
  using C = float;
  static constexpr bool exact = false;
  static constexpr int V = 8;
  __device__ static inline C cv(T v) { return (float)v * F16_INV; }
  __device__ static inline T enc(C v) { return to_f16(v); }
  __device__ static inline C ld(const T* p) { return (float)*p * F16_INV; }
  __device__ static inline void st(T* p, C v) { *p = to_f16(v); }
  __device__ static inline void st_stream(T* p, C v) { __builtin_nontemporal_store(to_f16(v), p); }
  __device__ static inline C round(C v) { return (float)to_f16(v) * F16_INV; }  // = ld(st(v))
};

template <class S, int M>
__device__ inline void collide(typename S::C* f, const Coll& c) {
  if constexpr (S::exact)
    Exact::collide<M>(f, c.op, c.om, c.gx, c.gy);
  else
    Dev::collide<M>(f, c.opf, c.omf, c.gxf, c.gyf);
}

__device__ inline int wrapi(int v, int n) { return v < 0 ? v + n : (v >= n ? v - n : v); }

// Pull the nine populations of node (x, y) from their upstream neighbours in A.
template <class S>
__device__ inline void pull(const typename S::T* __restrict__ A, const Geom& g, int x, int y,
                            typename S::C* f) {
  const int xm = wrapi(x - 1, g.nx), xp = wrapi(x + 1, g.nx);
  const int ym = wrapi(y - 1, g.ny), yp = wrapi(y + 1, g.ny);
  const int64_t r0 = (int64_t)y * g.pitch, rm = (int64_t)ym * g.pitch, rp = (int64_t)yp * g.pitch;
  const int64_t P = g.plane;
  f[0] = S::ld(A + r0 + x);
  f[1] = S::ld(A + P + r0 + xm);
  f[2] = S::ld(A + 2 * P + rm + x);
  f[3] = S::ld(A + 3 * P + r0 + xp);
  f[4] = S::ld(A + 4 * P + rp + x);
  f[5] = S::ld(A + 5 * P + rm + xm);
  f[6] = S::ld(A + 6 * P + rm + xp);
  f[7] = S::ld(A + 7 * P + rp + xp);
  f[8] = S::ld(A + 8 * P + rp + xm);
}

// ------------------------------------------------------------------------------------------
// The fused step kernel.  One launch per time step; workgroups [0, nlb) handle the boundary
// nodes (and the outlet controller and force reduction), the rest stream-collide the
// interior.  Boundary workgroups have the lowest ids so they are dispatched first and finish
// under the shadow of the bulk.
// ------------------------------------------------------------------------------------------
constexpr int STEP_T = 256;

struct StepArgs {
  int nlb;        // boundary workgroups (the force workgroup included)
  int fblk;       // workgroup that reduces the previous step's force terms (-1: none)
  int xblocks;    // bulk tiles per row (fused kernel)
  double sqrt_cs2;
  double* fc_cur;        // force terms written this step (per force slot: x, y)
  const double* fc_prev; // force terms of the previous step
  // Per-step controls and force history live in device tables indexed by the step's place in
  // the current run chunk, k = *kbase + k_local, so a captured window of steps can be replayed.
  const StepCtl* ctls;   // controls of the chunk's steps
  const int64_t* kbase;  // device counter: first chunk step of this window
  int k_local;           // step within the window
  double* hist;          // group forces of the chunk's steps (ngroups x 2 each), or null
  const double* pv_cur;  // outward velocity of every pressure-boundary node, this step
  double* pv_next;       // the same for the next step (written by the boundary update)
};

// Outlet controller (src/sim.rs:443-455): the mean outward normal velocity over the
// pressure-boundary nodes, as a density offset.  FP64 storage: the reference's serial
// left-to-right sum in node creation order (it feeds back into every later step, so the
// order matters for bit equality).  Deviation storage: a fixed-order tree.  Every thread of
// the workgroup must call this; all get the result.
// Outward normal velocity n.u of one node from its (stored-precision) populations, with the
// arithmetic of src/sim.rs:432-439, 447-449.
template <class S>
__device__ inline double normal_velocity_of(const typename S::C* f, const Coll& c, int nx_, int ny_) {
  if constexpr (S::exact) {
    double rho, jx, jy;
    Exact::moments(f, rho, jx, jy);
    const double ux = jx / rho - 0.5 * c.gx, uy = jy / rho - 0.5 * c.gy;
    return (double)nx_ * ux + (double)ny_ * uy;
  } else {
    float drho, jx, jy;
    Dev::moments(f, drho, jx, jy);
    const double rho = 1.0 + (double)drho;
    const double ux = (double)jx / rho - 0.5 * c.gx, uy = (double)jy / rho - 0.5 * c.gy;
    return (double)nx_ * ux + (double)ny_ * uy;
  }
}

// The same from the populations in A (used once at the start of a run).
template <class S>
__global__ __launch_bounds__(256) void k_outlet_prime(const typename S::T* __restrict__ A, Geom g, Coll c,
                                                      DevTables t, double* __restrict__ pv) {
  const int k = blockIdx.x * 256 + (int)threadIdx.x;
  if (k >= t.np) return;
  const int64_t idx = (int64_t)t.py[k] * g.pitch + t.px[k];
  typename S::C f[Q];
#pragma unroll
  for (int i = 0; i < Q; ++i) f[i] = S::ld(A + i * g.plane + idx);
  pv[k] = normal_velocity_of<S>(f, c, t.pnx[k], t.pny[k]);
}

// Density offset of the outlet controller from the nodes' outward velocities `pv` (which the
// previous step's boundary update left behind): serial left-to-right sum for FP64 storage,
// fixed-order tree otherwise.  Every thread of the workgroup must call this.
template <class S>
__device__ inline double outlet_offset(const double* __restrict__ pv, int np, double target, double sqrt_cs2,
                                double* red) {
  double sum = 0.0;  // serial running sum (exact) or this thread's strided partial (tree)
  for (int base = 0; base < np; base += STEP_T) {
    const int k = base + (int)threadIdx.x;
    const double v = k < np ? pv[k] : 0.0;
    if constexpr (S::exact) {
      red[threadIdx.x] = v;
      __syncthreads();
      if (threadIdx.x == 0) {
        const int m = min(STEP_T, np - base);
        for (int i = 0; i < m; ++i) sum += red[i];
      }
      __syncthreads();
    } else {
      sum += v;
    }
  }
  if constexpr (!S::exact) {
    red[threadIdx.x] = sum;
    __syncthreads();
    for (int w = STEP_T / 2; w > 0; w >>= 1) {
      if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
      __syncthreads();
    }
    sum = red[0];
    __syncthreads();
  }
  if (threadIdx.x == 0) red[0] = (sum / (double)np - target) / sqrt_cs2;
  __syncthreads();
  const double drho = red[0];
  __syncthreads();
  return drho;
}

// Force groups: fixed-order FP64 tree.  Thread k sums slots k, k+T, k+2T, ... in order, then
// a fixed pairwise tree over the threads: the same association every step and every run.
__device__ inline void reduce_forces(const double* __restrict__ fc, const int32_t* __restrict__ gfirst, int ngroups,
                              double* __restrict__ out, double* red) {
  for (int grp = 0; grp < ngroups; ++grp) {
    const int s0 = gfirst[grp], s1 = gfirst[grp + 1];
    for (int comp = 0; comp < 2; ++comp) {
      double a = 0.0;
      for (int s = s0 + (int)threadIdx.x; s < s1; s += STEP_T) a += fc[2 * s + comp];
      red[threadIdx.x] = a;
      __syncthreads();
      for (int w = STEP_T / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
      }
      if (threadIdx.x == 0) out[2 * grp + comp] = red[0];
      __syncthreads();
    }
  }
}

// One boundary node (src/sim.rs:456-507): pull, replace the populations that would come out
// of non-fluid cells by the link rules in link order, collide, write B; each force link
// stores its momentum-exchange term.
template <class S, int M>
__device__ inline void boundary_node(const typename S::T* __restrict__ A, typename S::T* __restrict__ B, const Geom& g,
                              const Coll& c, const DevTables& t, const StepCtl& ctl, double drho, double sqrt_cs2,
                              int b, double* __restrict__ fc, double* __restrict__ pv_next) {
  using C = typename S::C;
  const int x = t.bx[b], y = t.by[b];
  const int64_t idx = (int64_t)y * g.pitch + x;
  C f[Q], own[Q];
  pull<S>(A, g, x, y, f);
  // This node's own populations in A (what leaves it along each direction), loaded together
  // with the pull so that no link waits on a load of its own.
#pragma unroll
  for (int k = 0; k < Q; ++k) own[k] = S::ld(A + k * g.plane + idx);
  // Local velocity for anti-bounce-back links, computed once (the reference recomputes the
  // same value per link).
  bool have_u = false;
  double ux = 0.0, uy = 0.0;
  const int l1 = t.bfirst[b + 1];
  for (int l = t.bfirst[b]; l < l1; ++l) {
    const int j = t.dir[l];
    const int i = opp(j);
    const C out = own[j];
    const double scale = ctl.vscale[t.vgroup[l]];
    C back;
    if (t.kind[l] == LINK_WALL) {
      const double q = t.q[l];
      if constexpr (S::exact) {
        const double corr = t.cu6[l] * scale;
        if (q >= 0.5) {
          const double k = 0.5 / q;
          back = k * (out - corr) + (1.0 - k) * own[i];
        } else if (t.lflags[l] & 2) {
          const int bxn = wrapi(x - cx(j), g.nx), byn = wrapi(y - cy(j), g.ny);
          back = 2.0 * q * out + (1.0 - 2.0 * q) * A[j * g.plane + (int64_t)byn * g.pitch + bxn] - corr;
        } else {
          back = out - corr;
        }
      } else {
        // The wall rule is affine with weights summing to one, so it holds unchanged for
        // deviations (the w_i terms cancel).
        const float corr = (float)(t.cu6[l] * scale);
        const float qf = (float)q;
        if (q >= 0.5) {
          const float k = (float)(0.5 / q);
          back = k * (out - corr) + (1.0f - k) * own[i];
        } else if (t.lflags[l] & 2) {
          const int bxn = wrapi(x - cx(j), g.nx), byn = wrapi(y - cy(j), g.ny);
          back = 2.0f * qf * out + (1.0f - 2.0f * qf) * S::ld(A + j * g.plane + (int64_t)byn * g.pitch + bxn) -
                 corr;
        } else {
          back = out - corr;
        }
      }
    } else {
      if (!have_u) {
        have_u = true;
        if constexpr (S::exact) {
          double rho, jx, jy;
          Exact::moments(own, rho, jx, jy);
          ux = jx / rho - 0.5 * c.gx;
          uy = jy / rho - 0.5 * c.gy;
        } else {
          float dr, jx, jy;
          Dev::moments(own, dr, jx, jy);
          const double rho = 1.0 + (double)dr;
          ux = (double)jx / rho - 0.5 * c.gx;
          uy = (double)jy / rho - 0.5 * c.gy;
        }
      }
      const double cu = 3.0 * ((double)cx(j) * ux + (double)cy(j) * uy);
      const double off = (t.lflags[l] & 1) ? ((double)t.lnx[l] * ux + (double)t.lny[l] * uy) / sqrt_cs2 : drho;
      if constexpr (S::exact) {
        const double rho_w = t.rho[l] + off;
        back = -out + 2.0 * wgt(j) * rho_w * (1.0 + 0.5 * cu * cu - 1.5 * (ux * ux + uy * uy));
      } else {
        // back - w = -d_out + 2 w (drho_w + rho_w (cu^2/2 - 1.5 u^2))
        const double drw = (t.rho[l] - 1.0) + off;
        const double rho_w = 1.0 + drw;
        back = (C)(-(double)out + 2.0 * wgt(j) * (drw + rho_w * (0.5 * cu * cu - 1.5 * (ux * ux + uy * uy))));
      }
    }
    f[i] = back;
    const int slot = t.fslot[l];
    if (slot >= 0) {
      const double wx = t.uwx[l] * scale, wy = t.uwy[l] * scale;
      double o = (double)out, bk = (double)back;
      if constexpr (!S::exact) {
        o += wgt(j);
        bk += wgt(j);
      }
      fc[2 * slot] = (double)cx(j) * (o + bk) - wx * (o - bk);
      fc[2 * slot + 1] = (double)cy(j) * (o + bk) - wy * (o - bk);
    }
  }
  collide<S, M>(f, c);
#pragma unroll
  for (int i = 0; i < Q; ++i) S::st(B + i * g.plane + idx, f[i]);
  if (b < t.npb) {  // pressure-boundary node b is outlet-controller node b: its velocity next step
    C r[Q];
#pragma unroll
    for (int i = 0; i < Q; ++i) r[i] = S::round(f[i]);
    pv_next[b] = normal_velocity_of<S>(r, c, t.pnx[b], t.pny[b]);
  }
}

// Workgroup `blk` of the boundary part of a step.
template <class S, int M>
__device__ inline void boundary_block(int blk, const typename S::T* __restrict__ A, typename S::T* __restrict__ B,
                                      const Geom& g, const Coll& c, const DevTables& t, const StepArgs& sa,
                                      double* red) {
  const int64_t k = *sa.kbase + sa.k_local;
  if (blk == sa.fblk) {
    // a workgroup of its own: the force terms of the previous step of this chunk go to its
    // history slot, off the path of the boundary nodes
    if (sa.hist && k > 0) reduce_forces(sa.fc_prev, t.gfirst, t.ngroups, sa.hist + (k - 1) * 2 * t.ngroups, red);
    return;
  }
  const StepCtl ctl = sa.ctls[k];
  double drho = 0.0;
  if (ctl.has_target && t.np > 0 && blk * STEP_T < t.npb) {
    drho = outlet_offset<S>(sa.pv_cur, t.np, ctl.target, sa.sqrt_cs2, red);
    if (blk == 0 && threadIdx.x == 0) *t.drho = drho;
  }
  const int b = blk * STEP_T + (int)threadIdx.x;
  if (b < t.nb) boundary_node<S, M>(A, B, g, c, t, ctl, drho, sa.sqrt_cs2, b, sa.fc_cur, sa.pv_next);
}

// Interior span of a row, {lo, hi}: its first and last interior node (hi < lo: none).  A row
// whose span also holds boundary or solid nodes is stored with lo as -lo - 1, and only such
// rows read the per-node flag; elsewhere the span alone says which nodes are interior.
__device__ inline bool row_span(const int2* __restrict__ rows, int y, int& lo, int& hi) {
  const int2 r = rows[y];
  lo = r.x < 0 ? -r.x - 1 : r.x;
  hi = r.y;
  return r.x < 0;
}

__device__ inline bool is_interior(const int2* __restrict__ rows, const uint8_t* __restrict__ flag, int64_t idx,
                                   int x, int y) {
  if (!rows) return flag[idx] == 0;
  int lo, hi;
  const bool mixed = row_span(rows, y, lo, hi);
  return x >= lo && x <= hi && (!mixed || flag[idx] == 0);
}

// One interior node of bulk tile `tile` (tiles are `width` consecutive nodes of one row).
// `rows` (optional) holds each row's interior span; without it every node's flag is read.
template <class S, int M>
__device__ inline void bulk_node(int tile, int width, int xblocks, const typename S::T* __restrict__ A,
                                 typename S::T* __restrict__ B, const uint8_t* __restrict__ flag,
                                 const int2* __restrict__ rows, const Geom& g, const Coll& c) {
  const int y = tile / xblocks;
  const int x = (tile - y * xblocks) * width + (int)threadIdx.x;
  if (x >= g.nx) return;
  const int64_t idx = (int64_t)y * g.pitch + x;
  if (!is_interior(rows, flag, idx, x, y)) return;
  typename S::C f[Q];
  pull<S>(A, g, x, y, f);
  collide<S, M>(f, c);
#pragma unroll
  for (int i = 0; i < Q; ++i) S::st_stream(B + i * g.plane + idx, f[i]);
}

// V consecutive interior nodes of one row per thread (the split-mode bulk).  Each population
// plane is read as one aligned vector per thread (16 bytes at the storage's default V), plus one element beyond the vector's
// left or right end for the directions that shift along x, and written as one vector.  Threads
// whose V nodes are not all interior, or whose neighbours would wrap around a periodic x edge,
// take the one-node path for each of their nodes; the arithmetic per node is the same in both.
template <int V>
struct FlagWord;
template <>
struct FlagWord<2> {
  using T = uint16_t;
};
template <>
struct FlagWord<4> {
  using T = uint32_t;
};
template <>
struct FlagWord<8> {
  using T = uint64_t;
};

template <class S, int M, int V>
__device__ inline void bulk_vec(int tile, int width, int xblocks, const typename S::T* __restrict__ A,
                                typename S::T* __restrict__ B, const uint8_t* __restrict__ flag,
                                const int2* __restrict__ rows, const Geom& g, const Coll& c) {
  using T = typename S::T;
  using C = typename S::C;
  using VT = T __attribute__((ext_vector_type(V)));
  const int y = tile / xblocks;
  const int x0 = ((tile - y * xblocks) * width + (int)threadIdx.x) * V;
  if (x0 >= g.nx) return;
  const int64_t r0 = (int64_t)y * g.pitch;
  bool all_interior;
  if (rows) {
    int lo, hi;
    const bool mixed = row_span(rows, y, lo, hi);
    if (x0 + V - 1 < lo || x0 > hi) return;
    all_interior = x0 >= lo && x0 + V - 1 <= hi && (!mixed || *(const typename FlagWord<V>::T*)(flag + r0 + x0) == 0);
  } else {
    all_interior = *(const typename FlagWord<V>::T*)(flag + r0 + x0) == 0;
  }
  if (all_interior && x0 >= 1 && x0 + V < g.nx) {
    const int ym = wrapi(y - 1, g.ny), yp = wrapi(y + 1, g.ny);
    const int64_t rm = (int64_t)ym * g.pitch + x0, rp = (int64_t)yp * g.pitch + x0, rc = r0 + x0;
    const int64_t P = g.plane;
    const VT v0 = *(const VT*)(A + rc);
    const VT v1 = *(const VT*)(A + P + rc);
    const VT v2 = *(const VT*)(A + 2 * P + rm);
    const VT v3 = *(const VT*)(A + 3 * P + rc);
    const VT v4 = *(const VT*)(A + 4 * P + rp);
    const VT v5 = *(const VT*)(A + 5 * P + rm);
    const VT v6 = *(const VT*)(A + 6 * P + rm);
    const VT v7 = *(const VT*)(A + 7 * P + rp);
    const VT v8 = *(const VT*)(A + 8 * P + rp);
    const T l1 = A[P + rc - 1], r3 = A[3 * P + rc + V], l5 = A[5 * P + rm - 1], r6 = A[6 * P + rm + V];
    const T r7 = A[7 * P + rp + V], l8 = A[8 * P + rp - 1];
    VT o[Q];
#pragma unroll
    for (int k = 0; k < V; ++k) {
      const int km = k > 0 ? k - 1 : 0, kp = k < V - 1 ? k + 1 : V - 1;
      C f[Q];
      f[0] = S::cv(v0[k]);
      f[1] = S::cv(k == 0 ? l1 : v1[km]);
      f[2] = S::cv(v2[k]);
      f[3] = S::cv(k == V - 1 ? r3 : v3[kp]);
      f[4] = S::cv(v4[k]);
      f[5] = S::cv(k == 0 ? l5 : v5[km]);
      f[6] = S::cv(k == V - 1 ? r6 : v6[kp]);
      f[7] = S::cv(k == V - 1 ? r7 : v7[kp]);
      f[8] = S::cv(k == 0 ? l8 : v8[km]);
      collide<S, M>(f, c);
#pragma unroll
      for (int q = 0; q < Q; ++q) o[q][k] = S::enc(f[q]);
    }
#pragma unroll
    for (int q = 0; q < Q; ++q) __builtin_nontemporal_store(o[q], (VT*)(B + q * P + rc));
    return;
  }
#pragma unroll 1
  for (int k = 0; k < V; ++k) {
    const int x = x0 + k;
    if (x >= g.nx || !is_interior(rows, flag, r0 + x, x, y)) continue;
    C f[Q];
    pull<S>(A, g, x, y, f);
    collide<S, M>(f, c);
#pragma unroll
    for (int i = 0; i < Q; ++i) S::st_stream(B + i * g.plane + r0 + x, f[i]);
  }
}

// One row segment of W nodes per workgroup of W threads, with the six populations that shift
// along x staged through LDS: every global load of the segment is then aligned to the segment
// (plus one element on each side for the halo), and each thread reads its x - 1 / x + 1
// neighbours from LDS.  Nodes at x = 0 or x = nx - 1 (interior only in x-periodic domains) pull
// from global memory with the wrap.  Same arithmetic per node as bulk_node.
template <class S, int M, int W>
__device__ inline void bulk_lds(int tile, int xblocks, const typename S::T* __restrict__ A,
                                typename S::T* __restrict__ B, const uint8_t* __restrict__ flag,
                                const int2* __restrict__ rows, const Geom& g, const Coll& c) {
  using T = typename S::T;
  using C = typename S::C;
  __shared__ T sh[6][W + 2];  // planes 1, 3, 5, 6, 7, 8; element k holds x = xs - 1 + k
  const int y = tile / xblocks;
  const int xs = (tile - y * xblocks) * W;
  const int tid = (int)threadIdx.x;
  const int x = xs + tid;
  int lo = 0, hi = g.nx - 1;
  bool mixed = true;
  if (rows) {
    mixed = row_span(rows, y, lo, hi);
    if (xs > hi || xs + W - 1 < lo) return;  // the same for the whole workgroup
  }
  const int ym = wrapi(y - 1, g.ny), yp = wrapi(y + 1, g.ny);
  const int64_t r0 = (int64_t)y * g.pitch, rm = (int64_t)ym * g.pitch, rp = (int64_t)yp * g.pitch;
  const int64_t P = g.plane;
  const int64_t off[6] = {P + r0, 3 * P + r0, 5 * P + rm, 6 * P + rm, 7 * P + rp, 8 * P + rp};
  const bool in_row = x < g.nx;
  T d0 = T(0), d2 = T(0), d4 = T(0);
  if (in_row) {
    d0 = A[r0 + x];
    d2 = A[2 * P + rm + x];
    d4 = A[4 * P + rp + x];
  }
#pragma unroll
  for (int k = 0; k < 6; ++k) sh[k][tid + 1] = in_row ? A[off[k] + x] : T(0);
  if (tid == 0) {
    const int xl = xs > 0 ? xs - 1 : 0;  // x = 0 itself pulls through the global path
#pragma unroll
    for (int k = 0; k < 6; ++k) sh[k][0] = A[off[k] + xl];
  }
  if (tid == W - 1) {
    const int xr = xs + W < g.nx ? xs + W : g.nx - 1;  // likewise x = nx - 1
#pragma unroll
    for (int k = 0; k < 6; ++k) sh[k][W + 1] = A[off[k] + xr];
  }
  __syncthreads();
  if (!in_row || x < lo || x > hi || (mixed && flag[r0 + x] != 0)) return;
  C f[Q];
  if (x == 0 || x == g.nx - 1) {
    pull<S>(A, g, x, y, f);
  } else {
    f[0] = S::cv(d0);
    f[1] = S::cv(sh[0][tid]);
    f[2] = S::cv(d2);
    f[3] = S::cv(sh[1][tid + 2]);
    f[4] = S::cv(d4);
    f[5] = S::cv(sh[2][tid]);
    f[6] = S::cv(sh[3][tid + 2]);
    f[7] = S::cv(sh[4][tid + 2]);
    f[8] = S::cv(sh[5][tid]);
  }
  collide<S, M>(f, c);
#pragma unroll
  for (int i = 0; i < Q; ++i) S::st_stream(B + i * P + r0 + x, f[i]);
}

// ------------------------------------------------------------------------------------------
// Two time steps in one pass (temporal blocking of the bulk)
// ------------------------------------------------------------------------------------------
// One wave sweeps a strip of 64 columns upward through a band of rows.  Lane l owns column
// x0 - 2 + l.  At every row it loads that row's nine populations from A (time t), computes
// time t+1 one row below, and time t+2 two rows below; the three rows each level needs stay in
// registers, and the x - 1 / x + 1 neighbours come from the adjacent lanes.  Lanes 1..62 hold
// valid t+1 values, lanes 2..61 valid t+2 values, so a wave writes up to 60 columns.  Every
// node the wave computes at t+1 (its output plus a one-node margin) is interior, so no link
// rule applies and the arithmetic per node is the single-step bulk's: t+1 values are rounded
// through the storage type exactly as a store and reload would, and the result is bit-identical
// to two single steps.  Nodes whose t+1 value a single-step neighbour will pull (the task's
// "ring") are also written to B.
constexpr int TB_LANES = 64;
constexpr int TB_OUT = TB_LANES - 4;
constexpr int TB_DEPTH = 3;  // rows of loads in flight per wave

// Task of one wave: x = first output column, y = first output row, z = output rows,
// w = output columns (<= TB_OUT) | edge bits << 8.  Ring nodes (covered nodes whose t+1 value
// a single-step neighbour pulls) lie on the task's edges, since every other output node has all
// eight neighbours inside the task; edge bit 1 / 2 / 4 / 8 = the first row / last row / first
// column / last column holds some, and the whole edge then writes its t+1 values to B (values a
// single step would also produce, so the extra ones are harmless).
template <class C>
__device__ inline C from_left(C v) {  // lane l receives lane l - 1's value
  return __shfl_up(v, 1, TB_LANES);
}
template <class C>
__device__ inline C from_right(C v) {  // lane l receives lane l + 1's value
  return __shfl_down(v, 1, TB_LANES);
}

// A copy of v in a register of its own.  The unshifted populations of a row are carried to the
// next row; without an explicit copy the compiler keeps them in the load's register and moves
// the reload of that buffer elsewhere instead, which makes it wait for loads still in flight.
__device__ inline float fresh(float v) {
  float r;
  asm volatile("v_mov_b32 %0, %1" : "=&v"(r) : "v"(v));
  return r;
}
__device__ inline double fresh(double v) {
  double r;
  asm volatile("v_mov_b64 %0, %1" : "=&v"(r) : "v"(v));
  return r;
}

// Element `e` of plane i of a population array, from a wave-uniform base (the same for every
// lane).  The plane's base goes through a scalar move the optimiser cannot look into, so the
// address stays (scalar base) + (32-bit lane offset) instead of nine 64-bit vector addresses.
template <class T>
using gptr = __attribute__((address_space(1))) T*;  // global memory
template <class T>
__device__ inline gptr<T> plane_at(T* base, int64_t plane, int i, int e) {
  T* pb = base + i * plane;
  asm volatile("s_mov_b64 %0, %1" : "=s"(pb) : "s"(pb));
  return (gptr<T>)((__attribute__((address_space(1))) char*)pb + (uint32_t)e * (uint32_t)sizeof(T));
}

template <class S, int M>
__device__ inline void two_step_strip(const int4 task, const typename S::T* __restrict__ A,
                                      typename S::T* __restrict__ B, typename S::T* __restrict__ Cn,
                                      const Geom& g, const Coll& c) {
  using T = typename S::T;
  using C = typename S::C;
  const int lane = (int)(threadIdx.x & (TB_LANES - 1));
  const int x0 = task.x, y0 = task.y, rows = task.z, w = task.w & 0xff;
  const int edges = (task.w >> 8) & 15;
  // the workgroup's four tasks cover the same rows: its waves sweep in step, one barrier a row
  const bool lockstep = (task.w >> 12) & 1;
  const int xr = x0 - 2 + lane;                    // this lane's column
  const int xl = min(max(xr, 0), g.nx - 1);        // clamped for loads (lanes past the strip)
  const bool out_lane = lane >= 2 && lane < 2 + w;
  const int64_t P = g.plane;
  const int pitch = g.pitch;
  // Every lane reads the same row: the address of (xl, row) in plane i is the wave-uniform
  // A + i*P + row*pitch + (x0 - 2), plus the lane's constant offset xl - (x0 - 2).
  const int lane_off = xl - (x0 - 2);
  const int64_t row0 = (int64_t)(y0 - 2) * pitch + (x0 - 2);  // uniform part of row y0 - 2
  const int iters = rows + 4;  // rows y0 - 2 .. y0 + rows + 1 of time t
  // TB_DEPTH rows are loaded ahead: row k sits in buf[k % TB_DEPTH], and consuming it frees its
  // buffer for row k + TB_DEPTH (the rows past the last one reload the last one, so every load
  // is in bounds and unconditional).
  T buf[TB_DEPTH][Q];
#pragma unroll
  for (int b = 0; b < TB_DEPTH; ++b)
#pragma unroll
    for (int i = 0; i < Q; ++i) buf[b][i] = *plane_at(A + row0 + (int64_t)min(b, iters - 1) * pitch, P, i, lane_off);
  // level 0 (time t): row r-1's f0, f1, f3 and f2, f5, f6; row r-2's f2, f5, f6
  C p0 = C(0), p1 = C(0), p3 = C(0), qa2 = C(0), qa5 = C(0), qa6 = C(0), qb2 = C(0), qb5 = C(0), qb6 = C(0);
  // level 1 (time t+1): the same for the rows of t+1 values
  C gp0 = C(0), gp1 = C(0), gp3 = C(0), ga2 = C(0), ga5 = C(0), ga6 = C(0), gb2 = C(0), gb5 = C(0), gb6 = C(0);
  // lanes of the first / last output column that write their t+1 values to B on every row
  const bool ring_col = out_lane && (((edges & 4) && lane == 2) || ((edges & 8) && lane == 1 + w));
  // Iteration k consumes row r = y0 - 2 + k: first the row that arrived, then the stores of the
  // previous iteration's results and the load of row k + 2, then the compute.  The memory
  // counter that the wait for a load uses also counts stores, and the compiler waits for all
  // of them whenever the accesses in flight differ between paths; so the loads are
  // unconditional and follow the stores, and each wait covers accesses issued at least one
  // whole iteration earlier.  The loop runs in groups of TB_DEPTH iterations; the iterations
  // past the last one store the last row (the first of them) and otherwise keep nothing.
  auto iteration = [&](const int k, T (&cur)[Q]) __attribute__((always_inline)) {
    if (lockstep) __syncthreads();
    const C f0 = fresh(S::cv(cur[0])), f2 = fresh(S::cv(cur[2])), f4 = fresh(S::cv(cur[4]));
    const C f1 = from_left(S::cv(cur[1])), f5 = from_left(S::cv(cur[5])), f8 = from_left(S::cv(cur[8]));
    const C f3 = from_right(S::cv(cur[3])), f6 = from_right(S::cv(cur[6])), f7 = from_right(S::cv(cur[7]));
    __builtin_amdgcn_sched_barrier(0);
    {
      const T* rowp = A + row0 + (int64_t)min(k + TB_DEPTH, iters - 1) * pitch;
#pragma unroll
      for (int i = 0; i < Q; ++i) cur[i] = *plane_at(rowp, P, i, lane_off);
    }
    __builtin_amdgcn_sched_barrier(0);
    const int s = y0 - 3 + k;
    const bool ring_node = out_lane && s >= y0 && s < y0 + rows &&
                           (ring_col || ((edges & 1) && s == y0) || ((edges & 2) && s == y0 + rows - 1));
    if (k >= 2) {
      // time t+1 at row s = r - 1
      C e[Q] = {p0, p1, qb2, p3, f4, qb5, qb6, f7, f8};
      collide<S, M>(e, c);
#pragma unroll
      for (int i = 0; i < Q; ++i) e[i] = S::round(e[i]);
      if (ring_node) {
        T* rowp = B + (int64_t)s * pitch + (x0 - 2);
#pragma unroll
        for (int i = 0; i < Q; ++i) *plane_at(rowp, P, i, lane) = S::enc(e[i]);
      }
      const C e1 = from_left(e[1]), e5 = from_left(e[5]), e8 = from_left(e[8]);
      const C e3 = from_right(e[3]), e6 = from_right(e[6]), e7 = from_right(e[7]);
      if (k >= 4 && k < iters) {
        // time t+2 at row s - 1
        C h[Q] = {gp0, gp1, gb2, gp3, e[4], gb5, gb6, e7, e8};
        collide<S, M>(h, c);
        if (out_lane) {
          T* rowp = Cn + (int64_t)(s - 1) * pitch + (x0 - 2);
#pragma unroll
          for (int i = 0; i < Q; ++i) *plane_at(rowp, P, i, lane) = S::enc(h[i]);
        }
      }
      gb2 = ga2;
      gb5 = ga5;
      gb6 = ga6;
      ga2 = e[2];
      ga5 = e5;
      ga6 = e6;
      gp0 = e[0];
      gp1 = e1;
      gp3 = e3;
    }
    qb2 = qa2;
    qb5 = qa5;
    qb6 = qa6;
    qa2 = f2;
    qa5 = f5;
    qa6 = f6;
    p0 = f0;
    p1 = f1;
    p3 = f3;
  };
#pragma unroll 1
  for (int k = 0; k < iters; k += TB_DEPTH) {
#pragma unroll
    for (int b = 0; b < TB_DEPTH; ++b) iteration(k + b, buf[b]);
  }
}


// The single-step update of an interior node from a list (packed y << 16 | x), for the
// interior nodes that no two-step task covers.
template <class S, int M>
__device__ inline void listed_node(int k, int n, const uint32_t* __restrict__ nodes,
                                   const typename S::T* __restrict__ A, typename S::T* __restrict__ B,
                                   const Geom& g, const Coll& c) {
  if (k >= n) return;
  const uint32_t v = nodes[k];
  const int x = (int)(v & 0xffffu), y = (int)(v >> 16);
  typename S::C f[Q];
  pull<S>(A, g, x, y, f);
  collide<S, M>(f, c);
  const int64_t idx = (int64_t)y * g.pitch + x;
#pragma unroll
  for (int i = 0; i < Q; ++i) S::st_stream(B + i * g.plane + idx, f[i]);
}

}  // namespace vl
