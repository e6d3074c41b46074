// Device set-up of the cases of a batch: cell kinds, node flags and the boundary tables of
// every channel case whose body is a polygon or a circle, built by GPU kernels instead of the
// host builder (the reference's Builder::fill_rect / fill_shape and Sim::from_builder with
// cut_fraction, src/sim.rs:111-132, 219-354, src/geometry.rs:16-49, 268-279).  The tables come
// out in the host builder's order and packing, byte for byte: boundary nodes in row-major
// creation order with the outlet-controller nodes first, links in direction order, force slots
// in link order.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <functional>
#include <vector>

#include "setup.h"

namespace vl {

// One case as the set-up kernels see it.
struct SetupCase {
  int32_t nx, ny, pitch, inflow;  // inflow: VL_INFLOW_*
  int64_t cell0;                  // first cell of the case in the pitched kind / flag arrays
  int64_t row0;                   // first row of the case among all cases' rows
  double u, h;                    // inflow speed (mean speed of the parabola), channel height ny - 2
  int32_t has_kick, body;         // body: ShapeDesc::POLYGON or CIRCLE
  int32_t nverts, pad_;
  int64_t vert0;                  // first vertex (x, y pairs) in the vertex pool
  double kcx, kcy;                // centre of the body's spin during the kick
  double bcx, bcy, br;            // circle
  // packing offsets of the case's tables (filled in after the count pass)
  int64_t o_b, o_bf, o_l, o_p;
  int32_t npb, nl_needs;          // outlet-controller nodes and their links (from the count pass)
};

// Per-case totals of the count pass.
struct SetupCounts {
  int64_t nb, nl, nslots, npb, nl_needs, fluid;
  int32_t err;  // 1: a fluid cell touches the domain edge
  int32_t pad_;
};

// The packed device tables every case's DevTables points into.
struct SetupTables {
  int32_t *bx, *by, *bfirst, *fslot, *px, *py;
  uint8_t *dir, *kind, *vgroup, *lflags;
  int8_t *lnx, *lny, *pnx, *pny;
  double *q, *uwx, *uwy, *cu6, *rho;
};

// Allocates n bytes of device memory owned by the caller's batch (freed with it).
using SetupAlloc = std::function<vl_status(void**, size_t)>;

// Paints every case's cells, writes the pitched flags (`dflags`, total_cells bytes, already
// allocated) and counts each case's nodes, links, force slots, outlet-controller nodes and
// fluid cells; then allocates the packed tables through `alloc` at the host builder's offsets
// and fills them.  `verts`: the polygon vertices of all cases (x, y pairs).
vl_status device_setup(std::vector<SetupCase>& cases, const std::vector<double>& verts, uint8_t* dflags,
                       int64_t total_cells, hipStream_t stream, const SetupAlloc& alloc,
                       std::vector<SetupCounts>& counts, SetupTables& tables);

}  // namespace vl
