// Device code of one D2Q9 time step, shared by the single-domain solver (sim.hip) and the
// batched solver (batch.hip) so that both run exactly the same arithmetic per node.
#pragma once

#include <hip/hip_runtime.h>

#include <type_traits>

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
  __device__ static inline C quantize(C v) { return (float)to_f16(v) * F16_INV; }  // = ld(st(v))
  // Mass-conserving storage: the moving populations are rounded to the storage grid and their
  // rounding errors added to the rest population before it is rounded itself, so the stored
  // deviations sum to the collided mass up to the rest population's one rounding.  The rest
  // deviation is small (about w_0 (drho - 1.5 u^2)), so its FP16 step is far finer than the moving
  // ones'.  Plain rounding of all nine biases the density the same way at every node of a steady
  // flow, step after step: on a 4096 x 2048 circle wake it raised the mean drag by 5 % against
  // FP64 (0.7 % with this rule).  Carrying the momentum errors into populations 1 and 2 as well
  // made that wake 29 % off: the carried errors excite their non-hydrodynamic modes.  The values
  // come out exactly representable, so storing them rounds nothing more.
  __device__ static inline void conserve(C* f) {
    C r = 0.0f;
#pragma unroll
    for (int i = 1; i < Q; ++i) {
      const C q = quantize(f[i]);
      r += f[i] - q;
      f[i] = q;
    }
    f[0] = quantize(f[0] + r);
  }
};

// Collide.  The results are exactly representable in the storage type (FP16 storage rounds them
// here, carrying the moving populations' rounding into the rest population), so a value kept in
// registers equals what a store and a reload would give.
template <class S, int M>
__device__ inline void collide(typename S::C* f, const Coll& c) {
  if constexpr (S::exact) {
    Exact::collide<M>(f, c.op, c.om, c.gx, c.gy);
  } else {
    Dev::collide<M>(f, c.opf, c.omf, c.gxf, c.gyf);
    if constexpr (std::is_same<S, StoreF16>::value) S::conserve(f);
  }
}

__device__ inline int wrapi(int v, int n) { return v < 0 ? v + n : (v >= n ? v - n : v); }

// Where a node's populations live: `at(i, x, y)` is the address of population i of node (x, y).
// MainView: a population array in the pitched layout, as every single-domain kernel reads and
// writes it.  (The batch also sends the columns beside an inlet and an outlet to a side array,
// through a view of its own with the same interface.)
template <class T>
struct MainView {
  T* base;
  int64_t plane;
  int pitch;
  __device__ inline T* at(int i, int x, int y) const { return base + i * plane + ((int64_t)y * pitch + x); }
};

// Pull the nine populations of node (x, y) from their upstream neighbours through a view.
template <class S, class V>
__device__ inline void pull_v(const V& v, const Geom& g, int x, int y, typename S::C* f) {
  const int xm = wrapi(x - 1, g.nx), xp = wrapi(x + 1, g.nx);
  const int ym = wrapi(y - 1, g.ny), yp = wrapi(y + 1, g.ny);
  f[0] = S::ld(v.at(0, x, y));
  f[1] = S::ld(v.at(1, xm, y));
  f[2] = S::ld(v.at(2, x, ym));
  f[3] = S::ld(v.at(3, xp, y));
  f[4] = S::ld(v.at(4, x, yp));
  f[5] = S::ld(v.at(5, xm, ym));
  f[6] = S::ld(v.at(6, xp, ym));
  f[7] = S::ld(v.at(7, xp, yp));
  f[8] = S::ld(v.at(8, xm, yp));
}

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
// fixed-order tree otherwise.  Every thread of the workgroup must call this.  NT: threads of the
// workgroup (at least STEP_T); the first STEP_T take the sums, so the association is the same
// for every NT.
template <class S, int NT = STEP_T>
__device__ inline double outlet_offset(const double* __restrict__ pv, int np, double target, double sqrt_cs2,
                                double* red) {
  static_assert(NT >= STEP_T, "workgroup narrower than the reduction");
  const bool part = NT == STEP_T || (int)threadIdx.x < STEP_T;  // takes part in the sums
  double sum = 0.0;  // serial running sum (exact) or this thread's strided partial (tree)
  if constexpr (!S::exact) {
    // this thread's strided partial: the loads are independent, so eight are issued at a time
    // (the addition order stays slot k, k + T, k + 2T, ...)
    if (part) {
#pragma unroll 8
      for (int k = (int)threadIdx.x; k < np; k += STEP_T) sum += pv[k];
    }
  }
  for (int base = 0; S::exact && base < np; base += STEP_T) {
    const int k = base + (int)threadIdx.x;
    const double v = k < np ? pv[k] : 0.0;
    if constexpr (S::exact) {
      if (part) red[threadIdx.x] = v;
      __syncthreads();
      if (threadIdx.x == 0) {
        const int m = min(STEP_T, np - base);
        for (int i = 0; i < m; ++i) sum += red[i];
      }
      __syncthreads();
    }
  }
  if constexpr (!S::exact) {
    if (part) red[threadIdx.x] = sum;
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
// a fixed pairwise tree over the threads: the same association every step and every run.  NT:
// threads of the workgroup (at least STEP_T); the first STEP_T take the sums.
template <int NT = STEP_T>
__device__ inline void reduce_forces(const double* __restrict__ fc, const int32_t* __restrict__ gfirst, int ngroups,
                              double* __restrict__ out, double* red) {
  static_assert(NT >= STEP_T, "workgroup narrower than the reduction");
  const bool part = NT == STEP_T || (int)threadIdx.x < STEP_T;
  for (int grp = 0; grp < ngroups; ++grp) {
    const int s0 = gfirst[grp], s1 = gfirst[grp + 1];
    for (int comp = 0; comp < 2; ++comp) {
      double a = 0.0;
      if (part)
        for (int s = s0 + (int)threadIdx.x; s < s1; s += STEP_T) a += fc[2 * s + comp];
      if (part) red[threadIdx.x] = a;
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

// The links of boundary nodes in slot-major order: link k of boundary node b (its k-th link in
// link order) at [k * stride + b].  The small fields go into one word: direction (bits 0-3), kind
// (4-7), velocity group (8-11), flags (12-15), normal + 1 (16-17 x, 18-19 y) and, in slot 0, the
// node's link count (24-31).  The lanes of a wave hold consecutive boundary nodes (or consecutive
// rows of a column of them), so each field loads along a line, where the link tables' columns
// give every lane a line of its own.  A fluid node has at most one link per moving direction.
constexpr int LINK_SLOTS = 8;
struct LinkSlots {
  const uint32_t* meta = nullptr;
  const int32_t* fslot = nullptr;
  const double *q = nullptr, *cu6 = nullptr, *rho = nullptr, *uwx = nullptr, *uwy = nullptr;
  int64_t stride = 0;
};
__host__ __device__ inline uint32_t pack_link_meta(int dir, int kind, int vgroup, int lflags, int lnx, int lny,
                                                   int count) {
  return (uint32_t)dir | (uint32_t)kind << 4 | (uint32_t)vgroup << 8 | (uint32_t)lflags << 12 |
         (uint32_t)(lnx + 1) << 16 | (uint32_t)(lny + 1) << 18 | (uint32_t)count << 24;
}

// The fields of one link (the per-link columns of the boundary tables).
struct LinkFields {
  int dir, kind, vgroup, lflags, lnx, lny, fslot;
  double q, cu6, rho, uwx, uwy;
  int count;  // from slot 0 of a slot-major node: its link count
  // link k of boundary node b from the slot-major copy
  __device__ static inline LinkFields slot(const LinkSlots& s, int k, int b) {
    const int64_t e = (int64_t)k * s.stride + b;
    const uint32_t m = s.meta[e];
    LinkFields L;
    L.dir = (int)(m & 15u);
    L.kind = (int)((m >> 4) & 15u);
    L.vgroup = (int)((m >> 8) & 15u);
    L.lflags = (int)((m >> 12) & 15u);
    L.lnx = (int)((m >> 16) & 3u) - 1;
    L.lny = (int)((m >> 18) & 3u) - 1;
    L.count = (int)(m >> 24);
    L.fslot = s.fslot[e];
    L.q = s.q[e];
    L.cu6 = s.cu6[e];
    L.rho = s.rho[e];
    L.uwx = s.uwx[e];
    L.uwy = s.uwy[e];
    return L;
  }
  __device__ static inline LinkFields load(const DevTables& t, int l) {
    LinkFields L;
    L.dir = t.dir[l];
    L.kind = t.kind[l];
    L.vgroup = t.vgroup[l];
    L.lflags = t.lflags[l];
    L.lnx = t.lnx[l];
    L.lny = t.lny[l];
    L.fslot = t.fslot[l];
    L.q = t.q[l];
    L.cu6 = t.cu6[l];
    L.rho = t.rho[l];
    L.uwx = t.uwx[l];
    L.uwy = t.uwy[l];
    L.count = 0;
    return L;
  }
};

// One boundary node (src/sim.rs:456-507): pull, replace the populations that would come out
// of non-fluid cells by the link rules in link order, collide, write; each force link stores
// its momentum-exchange term.  Reads go through the view rv (time t), writes through wv.
// SLOTS: the links come from the slot-major copy ls instead of the link tables' columns (the same
// fields in the same order, so the same bits).
template <class S, int M, bool SLOTS = false, class RV, class WV>
__device__ inline void boundary_node_v(const RV& rv, const WV& wv, const Geom& g, const Coll& c, const DevTables& t,
                                       const StepCtl& ctl, double drho, double sqrt_cs2, int b,
                                       double* __restrict__ fc, double* __restrict__ pv_next,
                                       const LinkSlots* ls = nullptr) {
  using C = typename S::C;
  const int x = t.bx[b], y = t.by[b];
  C f[Q], own[Q];
  pull_v<S>(rv, g, x, y, f);
  // This node's own populations at time t (what leaves it along each direction), loaded
  // together with the pull so that no link waits on a load of its own.
#pragma unroll
  for (int k = 0; k < Q; ++k) own[k] = S::ld(rv.at(k, x, y));
  // Local velocity for anti-bounce-back links, computed once (the reference recomputes the
  // same value per link).
  bool have_u = false;
  double ux = 0.0, uy = 0.0;
  // Every field of a link is loaded together, unconditionally, and the next link's fields while
  // this one is processed: one memory round trip per node's link list instead of several per
  // link.  Indices past the table end are clamped (those loads are never used).
  // With slots, slot 0 carries the link count and the next slot loaded is clamped to the last
  // link.
  int l0 = 0, n = 0, lmax = 0;
  LinkFields nxt;
  if constexpr (SLOTS) {
    nxt = LinkFields::slot(*ls, 0, b);
    n = nxt.count;
  } else {
    l0 = t.bfirst[b];
    n = t.bfirst[b + 1] - l0;
    lmax = max(t.nl - 1, 0);
    nxt = LinkFields::load(t, min(l0, lmax));
  }
  for (int li = 0; li < n; ++li) {
    const LinkFields L = nxt;
    if constexpr (SLOTS)
      nxt = LinkFields::slot(*ls, min(li + 1, n - 1), b);
    else
      nxt = LinkFields::load(t, min(l0 + li + 1, lmax));
    const int j = L.dir;
    const int i = opp(j);
    const C out = own[j];
    // a select, not an indexed load: an indexed register array goes to the stack
    const int vg = L.vgroup;
    const double scale = vg == 0 ? ctl.vscale[0] : (vg == 1 ? ctl.vscale[1] : (vg == 2 ? ctl.vscale[2] : ctl.vscale[3]));
    C back;
    if (L.kind == LINK_WALL) {
      const double q = L.q;
      if constexpr (S::exact) {
        const double corr = L.cu6 * scale;
        if (q >= 0.5) {
          const double k = 0.5 / q;
          back = k * (out - corr) + (1.0 - k) * own[i];
        } else if (L.lflags & 2) {
          const int bxn = wrapi(x - cx(j), g.nx), byn = wrapi(y - cy(j), g.ny);
          back = 2.0 * q * out + (1.0 - 2.0 * q) * S::ld(rv.at(j, bxn, byn)) - corr;
        } else {
          back = out - corr;
        }
      } else {
        // The wall rule is affine with weights summing to one, so it holds unchanged for
        // deviations (the w_i terms cancel).
        const float corr = (float)(L.cu6 * scale);
        const float qf = (float)q;
        if (q >= 0.5) {
          const float k = (float)(0.5 / q);
          back = k * (out - corr) + (1.0f - k) * own[i];
        } else if (L.lflags & 2) {
          const int bxn = wrapi(x - cx(j), g.nx), byn = wrapi(y - cy(j), g.ny);
          back = 2.0f * qf * out + (1.0f - 2.0f * qf) * S::ld(rv.at(j, bxn, byn)) - corr;
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
      const double off = (L.lflags & 1) ? ((double)L.lnx * ux + (double)L.lny * uy) / sqrt_cs2 : drho;
      if constexpr (S::exact) {
        const double rho_w = L.rho + off;
        back = -out + 2.0 * wgt(j) * rho_w * (1.0 + 0.5 * cu * cu - 1.5 * (ux * ux + uy * uy));
      } else {
        // back - w = -d_out + 2 w (drho_w + rho_w (cu^2/2 - 1.5 u^2))
        const double drw = (L.rho - 1.0) + off;
        const double rho_w = 1.0 + drw;
        back = (C)(-(double)out + 2.0 * wgt(j) * (drw + rho_w * (0.5 * cu * cu - 1.5 * (ux * ux + uy * uy))));
      }
    }
    f[i] = back;
    const int slot = L.fslot;
    if (slot >= 0) {
      const double wx = L.uwx * scale, wy = L.uwy * scale;
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
  for (int i = 0; i < Q; ++i) S::st(wv.at(i, x, y), f[i]);
  if (b < t.npb) {  // pressure-boundary node b is outlet-controller node b: its velocity next step
    pv_next[b] = normal_velocity_of<S>(f, c, t.pnx[b], t.pny[b]);  // f is storage-exact
  }
}

// The same on population arrays in the pitched layout: reads A (time t), writes B.
template <class S, int M>
__device__ inline void boundary_node(const typename S::T* __restrict__ A, typename S::T* __restrict__ B, const Geom& g,
                                     const Coll& c, const DevTables& t, const StepCtl& ctl, double drho,
                                     double sqrt_cs2, int b, double* __restrict__ fc, double* __restrict__ pv_next) {
  using T = typename S::T;
  boundary_node_v<S, M>(MainView<const T>{A, g.plane, g.pitch}, MainView<T>{B, g.plane, g.pitch}, g, c, t, ctl, drho,
                        sqrt_cs2, b, fc, pv_next);
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
// K time steps in one pass (temporal blocking of the bulk)
// ------------------------------------------------------------------------------------------
// One wave sweeps a strip of 64 V columns upward through a band of rows; lane l owns the V
// columns xs + V l .. xs + V l + V - 1, where xs = x0 - tb_halo(K, V).  At every row r it loads that row's
// nine populations from A (time t, level 0; one V-vector per plane), and each level L = 1..K
// computes time t+L at row r - L from the three most recent rows of level L - 1, which stay in
// registers; the x - 1 / x + 1 neighbours come from the lane's own registers or, at the lane's
// edge elements, from the adjacent lanes.  Level L holds valid values L columns in from each
// side, so a wave writes up to tb_out(K, V) columns.  Every node the wave computes below level K
// (its output plus a margin of K - 1) is interior, so no link rule applies and the arithmetic
// per node is the single-step bulk's: the collide returns values exactly representable in the
// storage type, so values below level K equal what a store and reload would give, and the
// result is bit-identical to K single steps.
constexpr int TB_LANES = 64;
constexpr int TB_MAX_K = 5;  // steps per pass, at most
// Columns a wave holds left of its first output column: K rounded up to a multiple of V, so that
// every lane's V columns start at a multiple of V (aligned vector loads and stores), and output
// columns a wave writes: the rest, less the K columns on the right, a multiple of V.
__host__ __device__ constexpr int tb_halo(int k, int v) { return (k + v - 1) / v * v; }
__host__ __device__ constexpr int tb_out(int k, int v) { return (TB_LANES * v - tb_halo(k, v) - k) / v * v; }

// Task of one wave: x = first output column, y = first output row, z = output rows,
// w = output columns (<= tb_out(K, V), a multiple of V) | edge bits << 8 | lockstep << 12.
// Ring nodes (covered nodes whose intermediate values single-step neighbours pull) lie on the
// task's edges, since every other output node has all eight neighbours inside the task; edge bit
// 1 / 2 / 4 / 8 = the first row / last row / first column / last column holds some, and the whole
// edge then writes its values of levels 1..K-1 to the intermediate arrays: the three populations
// that point out of the task across that edge, the ones its uncovered neighbours pull (values
// the single steps would also produce, so the extra ones are harmless).  Lockstep: the workgroup's tasks
// cover the same rows, and its waves sweep in step (one barrier a row), so that together they
// read and write whole lines.  Joined (bit 13): the workgroup's W waves sweep one strip of
// 64 V W columns together, wave j loading the 64 V columns from the strip's first plus 64 V j,
// and they hand each other the x-neighbours of their edge lanes through LDS at every level
// (one barrier a level), so no wave recomputes a halo of its own; bits 16..23 then hold the
// wave's first output column less its first loaded column.
template <class C>
__device__ inline C from_left(C v) {  // lane l receives lane l - 1's value
  return __shfl_up(v, 1, TB_LANES);
}
template <class C>
__device__ inline C from_right(C v) {  // lane l receives lane l + 1's value
  return __shfl_down(v, 1, TB_LANES);
}
// The x - 1 (left) / x + 1 (right) neighbour of each of a lane's V columns.
template <class C, int V>
__device__ inline void left_of(const C (&v)[V], C (&o)[V]) {
  o[0] = from_left(v[V - 1]);
#pragma unroll
  for (int j = 1; j < V; ++j) o[j] = v[j - 1];
}
template <class C, int V>
__device__ inline void right_of(const C (&v)[V], C (&o)[V]) {
  o[V - 1] = from_right(v[0]);
#pragma unroll
  for (int j = 0; j + 1 < V; ++j) o[j] = v[j + 1];
}

// A copy of v in a register of its own.  The unshifted populations of a row are carried to the
// next row; without an explicit copy the compiler keeps them in the load's register and moves
// the reload of that buffer elsewhere instead, which makes it wait for loads still in flight.
__device__ inline float fresh(float v) {
  float r;
  asm volatile("v_mov_b32 %0, %1" : "=&v"(r) : "v"(v));
  return r;
}
__device__ inline double fresh(double v) {  // two 32-bit moves (no register-pair alignment to meet)
  const uint64_t b = __builtin_bit_cast(uint64_t, v);
  const float lo = fresh(__builtin_bit_cast(float, (uint32_t)b));
  const float hi = fresh(__builtin_bit_cast(float, (uint32_t)(b >> 32)));
  return __builtin_bit_cast(double, (uint64_t)__builtin_bit_cast(uint32_t, hi) << 32 | __builtin_bit_cast(uint32_t, lo));
}

// Element `e` of plane i of a population array, from a wave-uniform base (the same for every
// lane) plus a 32-bit lane offset, as a global-memory pointer.
template <class T>
using gptr = __attribute__((address_space(1))) T*;  // global memory
template <class T>
__device__ inline gptr<T> plane_at(T* base, int64_t plane, int i, int e) {
  return (gptr<T>)((__attribute__((address_space(1))) char*)(base + i * plane) + (uint32_t)e * (uint32_t)sizeof(T));
}

// V consecutive stored values, loaded and stored as one vector.
template <class T, int V>
struct Pack {
  using type = T __attribute__((ext_vector_type(V)));
};
template <class T>
struct Pack<T, 1> {
  using type = T;
};
template <class T, int V>
__device__ inline T elem(const typename Pack<T, V>::type& v, int j) {
  if constexpr (V == 1)
    return v;
  else
    return v[j];
}

// Store of V consecutive values to plane i of a population array at a wave-uniform row base plus
// a per-lane byte offset, as a buffer store.  The plane's row base goes into a resource
// descriptor built by scalar instructions, so no vector instruction forms a 64-bit address (the
// pass kernels are bound by vector instruction issue).  Word 3 is the gfx9 raw-buffer format
// (32-bit data format); the record count spans 4 GiB from the row base, past any byte offset a
// lane uses.
template <class T, int V>
__device__ inline void store_plane(T* rowbase, int64_t plane, int i, uint32_t byte_off,
                                   typename Pack<T, V>::type v) {
  const __amdgpu_buffer_rsrc_t r =
      __builtin_amdgcn_make_buffer_rsrc((void*)(rowbase + i * plane), (short)0, -1, 0x00020000);
  constexpr int B = (int)sizeof(T) * V;
  if constexpr (B == 2) {
    __builtin_amdgcn_raw_buffer_store_b16(__builtin_bit_cast(unsigned short, v), r, byte_off, 0, 0);
  } else if constexpr (B == 4) {
    __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, v), r, byte_off, 0, 0);
  } else {
    static_assert(B == 8 || B == 16, "store width");
    using U = unsigned __attribute__((ext_vector_type(B / 4)));
    if constexpr (B == 8)
      __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(U, v), r, byte_off, 0, 0);
    else
      __builtin_amdgcn_raw_buffer_store_b128(__builtin_bit_cast(U, v), r, byte_off, 0, 0);
  }
}

// Output of a pass: level K goes to `out`, levels 1..K-1 of the ring nodes to `ring[L - 1]`.
template <class T>
struct PassOut {
  T* ring[TB_MAX_K - 1];
  T* out;
};

// The three most recent rows of one level, as the next level pulls them: row r-1's f0, f1, f3
// (x-shifted where the pull needs it) and f2, f5, f6, and row r-2's f2, f5, f6, per column.
template <class C, int V>
struct Rows3 {
  C p0[V], p1[V], p3[V], qa2[V], qa5[V], qa6[V], qb2[V], qb5[V], qb6[V];
  // row r of this level (shifted) becomes row r-1
  __device__ inline void push(const C (&f)[Q][V]) {
#pragma unroll
    for (int j = 0; j < V; ++j) {
      qb2[j] = qa2[j];
      qb5[j] = qa5[j];
      qb6[j] = qa6[j];
      qa2[j] = f[2][j];
      qa5[j] = f[5][j];
      qa6[j] = f[6][j];
      p0[j] = f[0][j];
      p1[j] = f[1][j];
      p3[j] = f[3][j];
    }
  }
};

// Edge values a joined wave hands its neighbours at one level: lane 63's x + 1 pulls (f1, f5,
// f8 of its last column) for the wave on its right, lane 0's x - 1 pulls (f3, f6, f7 of its
// first column) for the wave on its left.  Two sets, used in turn, so that a wave writing the
// next level's values never meets a neighbour still reading this level's.
constexpr int TB_XCH = 6;

// JOIN: the task is joined (bit 13); then xch holds 2 nw TB_XCH values of LDS, wid is the wave's
// index in the workgroup and nw the workgroup's waves.  The two forms are separate code paths
// (each with its own register allocation).
template <class S, int M, int V, int D, int K, bool JOIN = false>
__device__ __attribute__((always_inline)) inline void multi_step_strip(const int4 task,
                                                                       const typename S::T* __restrict__ A,
                                                                       const PassOut<typename S::T> po,
                                                                       const Geom& g, const Coll& c,
                                                                       typename S::C* xch = nullptr, int wid = 0,
                                                                       int nw = 1) {
  using T = typename S::T;
  using C = typename S::C;
  using PT = typename Pack<T, V>::type;
  static_assert(K >= 2 && K <= TB_MAX_K, "steps per pass");
  const int lane = (int)(threadIdx.x & (TB_LANES - 1));
  const int x0 = task.x, y0 = task.y, rows = task.z, w = task.w & 0xff;
  const int edges = (task.w >> 8) & 15;
  const bool lockstep = (task.w >> 12) & 1;
  const int xs = x0 - (JOIN ? (task.w >> 16) & 0xff : tb_halo(K, V));  // first column (a multiple of V)
  const int xc = xs + V * lane;                        // this lane's first column
  const int xl = min(xc, g.pitch - V);                 // clamped for loads (lanes past the row)
  const bool out_lane = xc >= x0 && xc + V <= x0 + w;  // all V columns are output columns
  const int64_t P = g.plane;
  const int pitch = g.pitch;
  // Every lane reads the same row: the address of (xl, row) in plane i is the wave-uniform
  // A + i*P + row*pitch + xs, plus the lane's constant offset xl - xs.
  const int lane_off = xl - xs;
  const int64_t row0 = (int64_t)(y0 - K) * pitch + xs;  // uniform part of row y0 - K
  const int iters = rows + 2 * K;                       // rows y0 - K .. y0 + rows + K - 1 of time t
  // D rows are loaded ahead: row k sits in buf[k % D], and consuming it frees its buffer for row
  // k + D (the rows past the last one reload the last one, so every load is in bounds and
  // unconditional).
  PT buf[D][Q];
#pragma unroll
  for (int b = 0; b < D; ++b)
#pragma unroll
    for (int i = 0; i < Q; ++i)
      buf[b][i] = *(gptr<const PT>)plane_at(A + row0 + (int64_t)min(b, iters - 1) * pitch, P, i, lane_off);
  Rows3<C, V> st[K];  // levels 0..K-1, as their next level pulls them
#pragma unroll
  for (int L = 0; L < K; ++L)
#pragma unroll
    for (int j = 0; j < V; ++j)
      st[L].p0[j] = st[L].p1[j] = st[L].p3[j] = st[L].qa2[j] = st[L].qa5[j] = st[L].qa6[j] = st[L].qb2[j] =
          st[L].qb5[j] = st[L].qb6[j] = C(0);
  // the first / last output column, when it writes its intermediate values on every row
  bool ring_l[V], ring_r[V];
#pragma unroll
  for (int j = 0; j < V; ++j) {
    ring_l[j] = out_lane && (edges & 4) && xc + j == x0;
    ring_r[j] = out_lane && (edges & 8) && xc + j == x0 + w - 1;
  }
  // Joined waves: the newest row's shifted populations (fr) take their edge lanes' neighbours
  // from the adjacent waves (u: the row before the shift).
  int xi = 0;  // exchanges so far, choosing the LDS set
  const bool take_l = lane == 0 && wid > 0, take_r = lane == TB_LANES - 1 && wid + 1 < nw;
  const int wl = wid > 0 ? wid - 1 : wid, wr = wid + 1 < nw ? wid + 1 : wid;  // the neighbours (or itself)
  auto edge_swap = [&](C (&fr)[Q][V], const C (&u1)[V], const C (&u5)[V], const C (&u8)[V], const C (&u3)[V],
                       const C (&u6)[V], const C (&u7)[V]) __attribute__((always_inline)) {
    C* set = xch + (xi & 1) * nw * TB_XCH;
    if (lane == TB_LANES - 1) {
      set[wid * TB_XCH + 0] = u1[V - 1];
      set[wid * TB_XCH + 1] = u5[V - 1];
      set[wid * TB_XCH + 2] = u8[V - 1];
    }
    if (lane == 0) {
      set[wid * TB_XCH + 3] = u3[0];
      set[wid * TB_XCH + 4] = u6[0];
      set[wid * TB_XCH + 5] = u7[0];
    }
    __syncthreads();
    // every lane reads the neighbours' values (one broadcast read each); the edge lanes take them
    const C l1 = set[wl * TB_XCH + 0], l5 = set[wl * TB_XCH + 1], l8 = set[wl * TB_XCH + 2];
    const C r3 = set[wr * TB_XCH + 3], r6 = set[wr * TB_XCH + 4], r7 = set[wr * TB_XCH + 5];
    fr[1][0] = take_l ? l1 : fr[1][0];
    fr[5][0] = take_l ? l5 : fr[5][0];
    fr[8][0] = take_l ? l8 : fr[8][0];
    fr[3][V - 1] = take_r ? r3 : fr[3][V - 1];
    fr[6][V - 1] = take_r ? r6 : fr[6][V - 1];
    fr[7][V - 1] = take_r ? r7 : fr[7][V - 1];
    ++xi;
  };
  // Iteration k consumes row r = y0 - K + k: first the row that arrived, then the load of row
  // k + D, then the compute and its stores.  The memory counter that the wait for a load uses
  // also counts stores, and the compiler waits for all of them whenever the accesses in flight
  // differ between paths; so the loads are unconditional, and a wait for a row leaves the D - 1
  // rows loaded after it in flight.  Level L runs from iteration 2L on (its three source rows
  // exist); the loop runs in groups of D iterations, and the iterations past the last one keep
  // nothing.
  auto iteration = [&](const int k, PT (&cur)[Q]) __attribute__((always_inline)) {
    if (!JOIN && lockstep) __syncthreads();
    C fr[Q][V];  // the newest row of the level being consumed, x-shifted as the pull needs it
    {
      C u1[V], u3[V], u5[V], u6[V], u7[V], u8[V];
#pragma unroll
      for (int j = 0; j < V; ++j) {
        fr[0][j] = fresh(S::cv(elem<T, V>(cur[0], j)));
        fr[2][j] = fresh(S::cv(elem<T, V>(cur[2], j)));
        fr[4][j] = fresh(S::cv(elem<T, V>(cur[4], j)));
        u1[j] = S::cv(elem<T, V>(cur[1], j));
        u3[j] = S::cv(elem<T, V>(cur[3], j));
        u5[j] = S::cv(elem<T, V>(cur[5], j));
        u6[j] = S::cv(elem<T, V>(cur[6], j));
        u7[j] = S::cv(elem<T, V>(cur[7], j));
        u8[j] = S::cv(elem<T, V>(cur[8], j));
      }
      left_of(u1, fr[1]);
      left_of(u5, fr[5]);
      left_of(u8, fr[8]);
      right_of(u3, fr[3]);
      right_of(u6, fr[6]);
      right_of(u7, fr[7]);
      if constexpr (JOIN) edge_swap(fr, u1, u5, u8, u3, u6, u7);
    }
    __builtin_amdgcn_sched_barrier(0);
    {
      const T* rowp = A + row0 + (int64_t)min(k + D, iters - 1) * pitch;
#pragma unroll
      for (int i = 0; i < Q; ++i) cur[i] = *(gptr<const PT>)plane_at(rowp, P, i, lane_off);
    }
    __builtin_amdgcn_sched_barrier(0);
    bool active = true;  // the level below produced a row this iteration
#pragma unroll
    for (int L = 1; L <= K; ++L) {
      if (!active) break;
      const bool run = k >= 2 * L && (L < K || k < iters);
      C e[Q][V];
      if (run) {
        const int s = y0 - K + k - L;  // the row level L computes
        // Ring nodes' intermediate values are read only by the single steps of their uncovered
        // neighbours, which are interior nodes (covered nodes lie K - 1 nodes inside an
        // all-interior footprint), and each pulls the populations that point at it: row y0
        // writes those with cy = -1 (its uncovered neighbours lie below), the last row cy = +1,
        // the first column cx = -1, the last column cx = +1.
        const bool ring_row = L < K && s >= y0 && s < y0 + rows;
        const bool ring_lo = ring_row && out_lane && (edges & 1) && s == y0;
        const bool ring_hi = ring_row && out_lane && (edges & 2) && s == y0 + rows - 1;
        PT outv[Q];
#pragma unroll
        for (int j = 0; j < V; ++j) {
          const Rows3<C, V>& b = st[L - 1];
          C f[Q] = {b.p0[j], b.p1[j], b.qb2[j], b.p3[j], fr[4][j], b.qb5[j], b.qb6[j], fr[7][j], fr[8][j]};
          collide<S, M>(f, c);
          if (L < K) {
#pragma unroll
            for (int i = 0; i < Q; ++i) e[i][j] = f[i];  // storage-exact: as stored and reloaded
            const bool rl = ring_row && ring_l[j], rr = ring_row && ring_r[j];
            if (ring_lo || ring_hi || rl || rr) {
              T* rowp = po.ring[L - 1] + (int64_t)s * pitch + xs;
#pragma unroll
              for (int i = 1; i < Q; ++i)
                if ((ring_lo && cy(i) < 0) || (ring_hi && cy(i) > 0) || (rl && cx(i) < 0) || (rr && cx(i) > 0))
                  store_plane<T, 1>(rowp, P, i, (uint32_t)(V * lane + j) * (uint32_t)sizeof(T), S::enc(e[i][j]));
            }
          } else {
#pragma unroll
            for (int i = 0; i < Q; ++i) {
              if constexpr (V == 1)
                outv[i] = S::enc(f[i]);
              else
                outv[i][j] = S::enc(f[i]);
            }
          }
        }
        if (L == K && out_lane) {
          T* rowp = po.out + (int64_t)s * pitch + xs;
#pragma unroll
          for (int i = 0; i < Q; ++i) store_plane<T, V>(rowp, P, i, (uint32_t)(V * lane) * (uint32_t)sizeof(T), outv[i]);
        }
      }
      st[L - 1].push(fr);
      active = run && L < K;
      if (active) {  // level L's row, x-shifted, is what level L + 1 consumes
#pragma unroll
        for (int j = 0; j < V; ++j) {
          fr[0][j] = e[0][j];
          fr[2][j] = e[2][j];
          fr[4][j] = e[4][j];
        }
        C t1[V], t3[V], t5[V], t6[V], t7[V], t8[V];
#pragma unroll
        for (int j = 0; j < V; ++j) {
          t1[j] = e[1][j];
          t3[j] = e[3][j];
          t5[j] = e[5][j];
          t6[j] = e[6][j];
          t7[j] = e[7][j];
          t8[j] = e[8][j];
        }
        left_of(t1, fr[1]);
        left_of(t5, fr[5]);
        left_of(t8, fr[8]);
        right_of(t3, fr[3]);
        right_of(t6, fr[6]);
        right_of(t7, fr[7]);
        if constexpr (JOIN) edge_swap(fr, t1, t5, t8, t3, t6, t7);
      }
    }
  };
#pragma unroll 1
  for (int k = 0; k < iters; k += D) {
#pragma unroll
    for (int b = 0; b < D; ++b) iteration(k + b, buf[b]);
  }
}

// The single-step update of an interior node from a list (packed y << 16 | x), for the
// interior nodes that no K-step task covers.  Plain stores, not non-temporal ones: listed nodes
// come a few to a row (beside inlets, outlets and walls), so their stores fill partial lines that
// the boundary nodes' stores of the same step complete in L2 (the 512-case batch ran 3 % faster,
// the 8192 x 4096 FP32 job 1.4 %).
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
  for (int i = 0; i < Q; ++i) S::st(B + i * g.plane + idx, f[i]);
}

// The single-step update of interior node (x, y) through views: reads rv (time t), writes wv.
template <class S, int M, class RV, class WV>
__device__ inline void interior_node_v(int x, int y, const RV& rv, const WV& wv, const Geom& g, const Coll& c) {
  typename S::C f[Q];
  pull_v<S>(rv, g, x, y, f);
  collide<S, M>(f, c);
#pragma unroll
  for (int i = 0; i < Q; ++i) S::st(wv.at(i, x, y), f[i]);
}

}  // namespace vl
