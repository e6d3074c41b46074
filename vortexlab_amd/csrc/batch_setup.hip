// Device set-up of a batch's channel cases (see batch_setup.h).
//
// Three passes over every case's cells, all cases in one launch each:
//  * k_paint: the cell kinds, in the channel builder's painting order (inlet column, outlet
//    column, side rows, then the body, whose fill_shape claims every cell whose centre is
//    strictly inside it: src/sim.rs:111-132);
//  * k_classify (one workgroup per row): each fluid cell's links (non-fluid neighbours), force
//    links (into the body) and whether it feeds the outlet controller (a link into the outlet's
//    pressure cells); the node flags; per-row totals.  Exclusive scans of the row totals give
//    every row's first node, link and force slot within its case;
//  * k_fill (one workgroup per row): every boundary node and its links in creation order, at the
//    positions the host builder gives them after moving the outlet-controller nodes first.  The
//    body links' q is the 40-step bisection of the body predicate (src/geometry.rs:268-279) in
//    FP64 with the host's operation order, and the wall velocities are the channel's closures
//    evaluated at the wall point, so every field equals the host's bit for bit.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <string>

#include "batch_setup.h"
#include "lbm_math.h"
#include "sim_internal.h"

namespace vl {
namespace {

constexpr uint8_t K_FLUID = 0, K_INLET = 1, K_OUTLET = 2, K_SIDE = 3, K_BODY = 4, K_PAD = 255;
constexpr int SET_T = 256;
constexpr int NQ = 6;  // per-row quantities: nodes, links, force links, outlet nodes, their links, fluid cells

// Largest c with starts[c] <= v (starts ascending over nc entries).
__device__ inline int find_case(const int64_t* __restrict__ starts, int nc, int64_t v) {
  int lo = 0, hi = nc - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (starts[mid] <= v)
      lo = mid;
    else
      hi = mid - 1;
  }
  return lo;
}

// Bounding box of the body (with one cell of margin), outside of which the predicate is false.
struct BodyBox {
  double x0, x1, y0, y1;
};

// The body predicate: src/geometry.rs:16-27 (circle) and :30-49 (even-odd polygon), the same
// FP64 expressions as the host's.
__device__ inline bool body_inside(const SetupCase& c, const double* __restrict__ verts, double x, double y) {
  if (c.body == ShapeDesc::CIRCLE) {
    const double dx = x - c.bcx, dy = y - c.bcy;
    return dx * dx + dy * dy < c.br * c.br;
  }
  const double* p = verts + 2 * c.vert0;
  bool in = false;
  for (int i = 0, j = c.nverts - 1; i < c.nverts; j = i++) {
    const double xi = p[2 * i], yi = p[2 * i + 1], xj = p[2 * j], yj = p[2 * j + 1];
    if ((yi > y) != (yj > y) && x < (xj - xi) * (y - yi) / (yj - yi) + xi) in = !in;
  }
  return in;
}

// src/geometry.rs:268-279: 40 halvings of the link, then clamped to [1e-9, 1].
__device__ inline double cut_fraction_dev(const SetupCase& c, const double* __restrict__ verts, double x, double y,
                                          double dx, double dy) {
  double lo = 0.0, hi = 1.0;
  for (int k = 0; k < 40; ++k) {
    const double mid = 0.5 * (lo + hi);
    if (body_inside(c, verts, x + mid * dx, y + mid * dy))
      hi = mid;
    else
      lo = mid;
  }
  const double v = 0.5 * (lo + hi);
  const double m = v < 1e-9 ? 1e-9 : v;
  return 1.0 < m ? 1.0 : m;
}

__global__ __launch_bounds__(SET_T) void k_paint(const SetupCase* __restrict__ cases, const BodyBox* __restrict__ box,
                                                 int nc, const int64_t* __restrict__ cell_starts,
                                                 const double* __restrict__ verts, uint8_t* __restrict__ kinds,
                                                 int64_t total) {
  for (int64_t e = (int64_t)blockIdx.x * SET_T + threadIdx.x; e < total; e += (int64_t)gridDim.x * SET_T) {
    const int ci = find_case(cell_starts, nc, e);
    const SetupCase& c = cases[ci];
    const int64_t local = e - c.cell0;
    const int y = (int)(local / c.pitch), x = (int)(local - (int64_t)y * c.pitch);
    uint8_t k = K_FLUID;
    if (x >= c.nx) {
      k = K_PAD;
    } else {
      if (x == 0) k = K_INLET;
      if (x == c.nx - 1) k = K_OUTLET;
      if ((y == 0 || y == c.ny - 1) && ((x >= 1 && x <= c.nx - 2) || c.inflow != VL_INFLOW_UNIFORM_OPEN)) k = K_SIDE;
      const double fx = (double)x, fy = (double)y;
      const BodyBox& b = box[ci];
      if (fx >= b.x0 && fx <= b.x1 && fy >= b.y0 && fy <= b.y1 && body_inside(c, verts, fx, fy)) k = K_BODY;
    }
    kinds[e] = k;
  }
}

// Links of fluid cell (x, y): non-fluid neighbours, those into the body, and whether one goes
// into the outlet's pressure cells.  `edge`: a neighbour lies outside the domain.
struct CellLinks {
  int links, body, needs;
};
__device__ inline CellLinks cell_links(const SetupCase& c, const uint8_t* __restrict__ kc, int x, int y, bool& edge) {
  CellLinks o{0, 0, 0};
#pragma unroll
  for (int j = 1; j < Q; ++j) {
    const int xn = x + cx(j), yn = y + cy(j);
    if (xn < 0 || xn >= c.nx || yn < 0 || yn >= c.ny) {
      edge = true;
      continue;
    }
    const uint8_t kn = kc[(int64_t)yn * c.pitch + xn];
    if (kn == K_FLUID) continue;
    ++o.links;
    if (kn == K_BODY) ++o.body;
    if (kn == K_OUTLET) o.needs = 1;
  }
  return o;
}

struct Sums {
  uint32_t v[NQ];
  __device__ Sums operator+(const Sums& o) const {
    Sums r;
#pragma unroll
    for (int q = 0; q < NQ; ++q) r.v[q] = v[q] + o.v[q];
    return r;
  }
};
struct SumOp {
  __device__ Sums operator()(const Sums& a, const Sums& b) const { return a + b; }
};

__device__ inline Sums cell_sums(const SetupCase& c, const uint8_t* __restrict__ kc, int x, int y, bool& edge) {
  Sums s{};
  if (x >= c.nx || kc[(int64_t)y * c.pitch + x] != K_FLUID) return s;
  const CellLinks l = cell_links(c, kc, x, y, edge);
  const uint32_t node = l.links > 0;
  s.v[0] = node;
  s.v[1] = (uint32_t)l.links;
  s.v[2] = (uint32_t)l.body;
  s.v[3] = node && l.needs;
  s.v[4] = node && l.needs ? (uint32_t)l.links : 0u;
  s.v[5] = 1;
  return s;
}

// Flags and per-row totals.  rowcnt: NQ arrays of nrows + 1 entries (the last one stays 0, so
// an exclusive scan of nrows + 1 entries ends in the total).
__global__ __launch_bounds__(SET_T) void k_classify(const SetupCase* __restrict__ cases, int nc,
                                                    const int64_t* __restrict__ row_starts,
                                                    const uint8_t* __restrict__ kinds, uint8_t* __restrict__ flags,
                                                    uint32_t* __restrict__ rowcnt, int64_t stride,
                                                    int32_t* __restrict__ err) {
  using BlockReduce = hipcub::BlockReduce<Sums, SET_T>;
  __shared__ typename BlockReduce::TempStorage tmp;
  const int64_t r = blockIdx.x;
  const int ci = find_case(row_starts, nc, r);
  const SetupCase& c = cases[ci];
  const int y = (int)(r - c.row0);
  const uint8_t* kc = kinds + c.cell0;
  Sums acc{};
  bool edge = false;
  for (int x = (int)threadIdx.x; x < c.pitch; x += SET_T) {
    const Sums s = cell_sums(c, kc, x, y, edge);
    acc = acc + s;
    flags[c.cell0 + (int64_t)y * c.pitch + x] = x >= c.nx || kc[(int64_t)y * c.pitch + x] != K_FLUID ? 2 : (s.v[0] ? 1 : 0);
  }
  if (edge) atomicOr(&err[ci], 1);
  const Sums tot = BlockReduce(tmp).Reduce(acc, SumOp());
  if (threadIdx.x == 0)
#pragma unroll
    for (int q = 0; q < NQ; ++q) rowcnt[q * stride + r] = tot.v[q];
}

// Per-case totals from the row scans (exclusive, nrows + 1 entries each).
__global__ void k_totals(const SetupCase* __restrict__ cases, int nc, const int64_t* __restrict__ row_starts,
                         const uint32_t* __restrict__ rowscan, int64_t stride, const int32_t* __restrict__ err,
                         SetupCounts* __restrict__ out) {
  const int ci = blockIdx.x * blockDim.x + threadIdx.x;
  if (ci >= nc) return;
  const int64_t r0 = row_starts[ci], r1 = row_starts[ci + 1];
  auto tot = [&](int q) { return (int64_t)(rowscan[q * stride + r1] - rowscan[q * stride + r0]); };
  SetupCounts o{};
  o.nb = tot(0);
  o.nl = tot(1);
  o.nslots = tot(2);
  o.npb = tot(3);
  o.nl_needs = tot(4);
  o.fluid = tot(5);
  o.err = err[ci];
  out[ci] = o;
}

// Wall velocity of a link into cell kind `kn` at the wall point (px, py): the channel's closures
// (src/cases/channel.rs:95-103, 122; inlet profile, moving side walls, body spin).  Returns false
// for a wall at rest.
__device__ inline bool wall_velocity(const SetupCase& c, uint8_t kn, double px, double py, double& ux, double& uy) {
  if (kn == K_INLET) {
    if (c.inflow == VL_INFLOW_PARABOLIC) {
      const double h = c.h;
      const double t = py - 0.5;
      const double lo = t < 0.0 ? 0.0 : t;  // std::max(y - 0.5, 0.0)
      const double s = h < lo ? h : lo;     // std::min(., h)
      ux = 6.0 * c.u * s * (h - s) / (h * h);
      uy = 0.0;
    } else {
      ux = c.u;
      uy = 0.0;
    }
    return true;
  }
  if (kn == K_SIDE) {
    if (c.inflow != VL_INFLOW_UNIFORM) return false;
    ux = c.u;
    uy = 0.0;
    return true;
  }
  if (kn == K_BODY && c.has_kick) {
    ux = -(py - c.kcy);
    uy = px - c.kcx;
    return true;
  }
  return false;
}

// Every boundary node of one row and its links.  Thread t takes a contiguous chunk of the row's
// cells, so a workgroup scan of the per-thread sums gives each node's place in creation order.
__global__ __launch_bounds__(SET_T) void k_fill(const SetupCase* __restrict__ cases, int nc,
                                                const int64_t* __restrict__ row_starts,
                                                const uint8_t* __restrict__ kinds, const double* __restrict__ verts,
                                                const uint32_t* __restrict__ rowscan, int64_t stride, SetupTables T) {
  using BlockScan = hipcub::BlockScan<Sums, SET_T>;
  __shared__ typename BlockScan::TempStorage tmp;
  const int64_t r = blockIdx.x;
  const int ci = find_case(row_starts, nc, r);
  const SetupCase& c = cases[ci];
  const int y = (int)(r - c.row0);
  const uint8_t* kc = kinds + c.cell0;
  const int chunk = (c.nx + SET_T - 1) / SET_T;
  const int xa = min((int)threadIdx.x * chunk, c.nx), xb = min(xa + chunk, c.nx);
  bool edge = false;
  Sums mine{};
  for (int x = xa; x < xb; ++x) mine = mine + cell_sums(c, kc, x, y, edge);
  Sums before;
  BlockScan(tmp).ExclusiveScan(mine, before, Sums{}, SumOp());
  // case-local counts of everything before this thread's first cell
  uint32_t B = before.v[0], L = before.v[1], BL = before.v[2], N = before.v[3], NL = before.v[4];
  B += rowscan[0 * stride + r] - rowscan[0 * stride + c.row0];
  L += rowscan[1 * stride + r] - rowscan[1 * stride + c.row0];
  BL += rowscan[2 * stride + r] - rowscan[2 * stride + c.row0];
  N += rowscan[3 * stride + r] - rowscan[3 * stride + c.row0];
  NL += rowscan[4 * stride + r] - rowscan[4 * stride + c.row0];
  for (int x = xa; x < xb; ++x) {
    const Sums s = cell_sums(c, kc, x, y, edge);
    if (!s.v[0]) continue;
    const bool needs = s.v[3] != 0;
    // outlet-controller nodes first (in creation order), then the others (in creation order)
    const int64_t p = needs ? (int64_t)N : (int64_t)c.npb + (B - N);
    const int64_t lo = needs ? (int64_t)NL : (int64_t)c.nl_needs + (L - NL);
    T.bx[c.o_b + p] = x;
    T.by[c.o_b + p] = y;
    T.bfirst[c.o_bf + p] = (int32_t)lo;
    int k = 0;
    uint32_t slot = BL;
    bool pushed = false;
#pragma unroll 1
    for (int j = 1; j < Q; ++j) {
      const int xn = x + cx(j), yn = y + cy(j);
      const uint8_t kn = kc[(int64_t)yn * c.pitch + xn];
      if (kn == K_FLUID) continue;
      const int64_t l = c.o_l + lo + k;
      ++k;
      const int xbk = x - cx(j), ybk = y - cy(j);
      const bool behind = xbk >= 0 && xbk < c.nx && ybk >= 0 && ybk < c.ny && kc[(int64_t)ybk * c.pitch + xbk] == K_FLUID;
      uint8_t lflags = behind ? 2 : 0, lkind = LINK_WALL, vgroup = 0;
      int8_t nxo = 0, nyo = 0;
      double q = 0.5, uwx = 0.0, uwy = 0.0, cu6 = 0.0, rho = 1.0;
      int32_t fslot = -1;
      const bool open_side = kn == K_SIDE && c.inflow == VL_INFLOW_UNIFORM_OPEN;
      if (kn == K_OUTLET || open_side) {
        lkind = LINK_PRESSURE;
        rho = 1.0;
        if (open_side) lflags |= 1;
        // outward normal along x if the cell beside the node in x is open too, else along y
        const int xs = x + cx(j);
        bool side_open = false;
        if (xs >= 0 && xs < c.nx) {
          const uint8_t ks = kc[(int64_t)y * c.pitch + xs];
          side_open = ks == K_OUTLET || (ks == K_SIDE && c.inflow == VL_INFLOW_UNIFORM_OPEN);
        }
        if (side_open)
          nxo = (int8_t)cx(j);
        else
          nyo = (int8_t)cy(j);
        if (!open_side && !pushed) {
          T.px[c.o_p + N] = x;
          T.py[c.o_p + N] = y;
          T.pnx[c.o_p + N] = nxo;
          T.pny[c.o_p + N] = nyo;
          pushed = true;
        }
      } else {
        vgroup = kn == K_INLET ? 1 : (kn == K_SIDE ? (c.inflow == VL_INFLOW_UNIFORM ? 1 : 0) : 2);
        if (kn == K_BODY) q = cut_fraction_dev(c, verts, (double)x, (double)y, (double)cx(j), (double)cy(j));
        double ux, uy;
        if (wall_velocity(c, kn, (double)x + q * (double)cx(j), (double)y + q * (double)cy(j), ux, uy)) {
          uwx = ux;
          uwy = uy;
          cu6 = 6.0 * wgt(j) * ((double)cx(j) * uwx + (double)cy(j) * uwy);
        }
        if (kn == K_BODY) fslot = (int32_t)slot++;
      }
      T.dir[l] = (uint8_t)j;
      T.kind[l] = lkind;
      T.vgroup[l] = vgroup;
      T.lflags[l] = lflags;
      T.lnx[l] = nxo;
      T.lny[l] = nyo;
      T.fslot[l] = fslot;
      T.q[l] = q;
      T.uwx[l] = uwx;
      T.uwy[l] = uwy;
      T.cu6[l] = cu6;
      T.rho[l] = rho;
    }
    B += 1;
    L += s.v[1];
    BL += s.v[2];
    N += s.v[3];
    NL += s.v[4];
  }
}

// The end marker of each case's node -> first-link table.
__global__ void k_close(const SetupCase* __restrict__ cases, const SetupCounts* __restrict__ counts, int nc,
                        int32_t* __restrict__ bfirst) {
  const int ci = blockIdx.x * blockDim.x + threadIdx.x;
  if (ci < nc) bfirst[cases[ci].o_bf + counts[ci].nb] = (int32_t)counts[ci].nl;
}

// Device buffers of the set-up, freed on every path out.
struct Scratch {
  std::vector<void*> p;
  template <class V>
  vl_status get(V*& out, size_t n) {
    void* q = nullptr;
    VL_HIP(hipMalloc(&q, std::max<size_t>(n, 1) * sizeof(V)));
    p.push_back(q);
    out = (V*)q;
    return VL_OK;
  }
  ~Scratch() {
    for (void* q : p) (void)hipFree(q);
  }
};

#define VL_TRYS(x)                \
  do {                            \
    vl_status st_ = (x);          \
    if (st_ != VL_OK) return st_; \
  } while (0)

template <class V>
vl_status put(Scratch& s, V*& d, const std::vector<V>& h, hipStream_t st) {
  VL_TRYS(s.get(d, h.size()));
  if (!h.empty()) VL_HIP(hipMemcpyAsync(d, h.data(), h.size() * sizeof(V), hipMemcpyHostToDevice, st));
  return VL_OK;
}

}  // namespace

vl_status device_setup(std::vector<SetupCase>& cases, const std::vector<double>& verts, uint8_t* dflags,
                       int64_t total_cells, hipStream_t stream, const SetupAlloc& alloc,
                       std::vector<SetupCounts>& counts, SetupTables& T) {
  const int nc = (int)cases.size();
  if (nc == 0) return set_error(VL_EINVAL, "device set-up of no cases");
  Scratch s;
  std::vector<int64_t> cell_starts(nc + 1), row_starts(nc + 1);
  std::vector<BodyBox> boxes(nc);
  int64_t rows = 0;
  for (int i = 0; i < nc; ++i) {
    SetupCase& c = cases[i];
    cell_starts[i] = c.cell0;
    c.row0 = rows;
    row_starts[i] = rows;
    rows += c.ny;
    BodyBox& b = boxes[i];
    if (c.body == ShapeDesc::CIRCLE) {
      b = {c.bcx - c.br - 1.0, c.bcx + c.br + 1.0, c.bcy - c.br - 1.0, c.bcy + c.br + 1.0};
    } else {
      b = {1e300, -1e300, 1e300, -1e300};
      for (int k = 0; k < c.nverts; ++k) {
        const double x = verts[(size_t)(2 * (c.vert0 + k))], y = verts[(size_t)(2 * (c.vert0 + k) + 1)];
        b.x0 = std::min(b.x0, x - 1.0);
        b.x1 = std::max(b.x1, x + 1.0);
        b.y0 = std::min(b.y0, y - 1.0);
        b.y1 = std::max(b.y1, y + 1.0);
      }
    }
  }
  cell_starts[nc] = total_cells;
  row_starts[nc] = rows;
  SetupCase* d_cases;
  BodyBox* d_box;
  int64_t *d_cs, *d_rs;
  double* d_verts;
  VL_TRYS(put(s, d_cases, cases, stream));
  VL_TRYS(put(s, d_box, boxes, stream));
  VL_TRYS(put(s, d_cs, cell_starts, stream));
  VL_TRYS(put(s, d_rs, row_starts, stream));
  VL_TRYS(put(s, d_verts, verts, stream));
  uint8_t* kinds;
  VL_TRYS(s.get(kinds, (size_t)total_cells));
  const int64_t stride = rows + 1;
  uint32_t *rowcnt, *rowscan;
  int32_t* d_err;
  VL_TRYS(s.get(rowcnt, (size_t)(NQ * stride)));
  VL_TRYS(s.get(rowscan, (size_t)(NQ * stride)));
  VL_TRYS(s.get(d_err, (size_t)nc));
  VL_HIP(hipMemsetAsync(rowcnt, 0, (size_t)(NQ * stride) * sizeof(uint32_t), stream));
  VL_HIP(hipMemsetAsync(d_err, 0, (size_t)nc * sizeof(int32_t), stream));
  const unsigned pgrid = (unsigned)std::min<int64_t>((total_cells + SET_T - 1) / SET_T, 1 << 20);
  hipLaunchKernelGGL(k_paint, dim3(pgrid), dim3(SET_T), 0, stream, d_cases, d_box, nc, d_cs, d_verts, kinds,
                     total_cells);
  hipLaunchKernelGGL(k_classify, dim3((unsigned)rows), dim3(SET_T), 0, stream, d_cases, nc, d_rs, kinds, dflags,
                     rowcnt, stride, d_err);
  VL_HIP(hipGetLastError());
  size_t tmp_bytes = 0;
  VL_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, tmp_bytes, rowcnt, rowscan, (int)stride, stream));
  char* tmp;
  VL_TRYS(s.get(tmp, tmp_bytes));
  for (int q = 0; q < NQ; ++q)
    VL_HIP(hipcub::DeviceScan::ExclusiveSum(tmp, tmp_bytes, rowcnt + q * stride, rowscan + q * stride, (int)stride,
                                            stream));
  SetupCounts* d_counts;
  VL_TRYS(s.get(d_counts, (size_t)nc));
  hipLaunchKernelGGL(k_totals, dim3((unsigned)((nc + 255) / 256)), dim3(256), 0, stream, d_cases, nc, d_rs, rowscan,
                     stride, d_err, d_counts);
  VL_HIP(hipGetLastError());
  counts.assign(nc, SetupCounts{});
  VL_HIP(hipMemcpyAsync(counts.data(), d_counts, (size_t)nc * sizeof(SetupCounts), hipMemcpyDeviceToHost, stream));
  VL_HIP(hipStreamSynchronize(stream));
  // packing offsets: the host builder's concatenation of the cases' tables
  int64_t nb = 0, nbf = 0, nl = 0, np = 0;
  for (int i = 0; i < nc; ++i) {
    const SetupCounts& k = counts[i];
    if (k.err) return set_error(VL_EINVAL, "case " + std::to_string(i) + ": a fluid cell touches a non-periodic edge");
    SetupCase& c = cases[i];
    c.o_b = nb;
    c.o_bf = nbf;
    c.o_l = nl;
    c.o_p = np;
    c.npb = (int32_t)k.npb;
    c.nl_needs = (int32_t)k.nl_needs;
    nb += k.nb;
    nbf += k.nb + 1;
    nl += k.nl;
    np += k.npb;
  }
  VL_HIP(hipMemcpyAsync(d_cases, cases.data(), (size_t)nc * sizeof(SetupCase), hipMemcpyHostToDevice, stream));
  auto A = [&](auto*& ptr, int64_t n) {
    return alloc((void**)&ptr, (size_t)std::max<int64_t>(n, 1) * sizeof(*ptr));
  };
  VL_TRYS(A(T.bx, nb));
  VL_TRYS(A(T.by, nb));
  VL_TRYS(A(T.bfirst, nbf));
  VL_TRYS(A(T.fslot, nl));
  VL_TRYS(A(T.px, np));
  VL_TRYS(A(T.py, np));
  VL_TRYS(A(T.dir, nl));
  VL_TRYS(A(T.kind, nl));
  VL_TRYS(A(T.vgroup, nl));
  VL_TRYS(A(T.lflags, nl));
  VL_TRYS(A(T.lnx, nl));
  VL_TRYS(A(T.lny, nl));
  VL_TRYS(A(T.pnx, np));
  VL_TRYS(A(T.pny, np));
  VL_TRYS(A(T.q, nl));
  VL_TRYS(A(T.uwx, nl));
  VL_TRYS(A(T.uwy, nl));
  VL_TRYS(A(T.cu6, nl));
  VL_TRYS(A(T.rho, nl));
  hipLaunchKernelGGL(k_fill, dim3((unsigned)rows), dim3(SET_T), 0, stream, d_cases, nc, d_rs, kinds, d_verts,
                     rowscan, stride, T);
  hipLaunchKernelGGL(k_close, dim3((unsigned)((nc + 255) / 256)), dim3(256), 0, stream, d_cases, d_counts, nc,
                     T.bfirst);
  VL_HIP(hipGetLastError());
  VL_HIP(hipStreamSynchronize(stream));
  return VL_OK;
}

}  // namespace vl
