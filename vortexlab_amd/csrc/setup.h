// Host-side set-up of a domain: shapes, the cell-painting builder and the boundary tables
// handed to the GPU.  This is the once-per-run part of the reference's Builder/from_builder
// (src/sim.rs:70-354) and geometry (src/geometry.rs); the step never calls back into it.
#pragma once

#include <array>
#include <cstdint>
#include <functional>
#include <memory>
#include <string>
#include <vector>

#include "../../include/vortexlab.h"

namespace vl {

// ---- error reporting (thread-local message behind vl_last_error) ------------------------
vl_status set_error(vl_status st, const std::string& msg);

// ---- shapes ------------------------------------------------------------------------------
// A shape the device set-up of a batch can evaluate itself: a polygon (even-odd rule over its
// vertices) or a circle; other shapes are evaluated on the host.
struct ShapeDesc {
  enum Kind { NONE = 0, POLYGON = 1, CIRCLE = 2 } kind = NONE;
  std::vector<std::array<double, 2>> pts;
  double cx = 0.0, cy = 0.0, r = 0.0;
};
struct Shape {
  virtual ~Shape() = default;
  virtual bool inside(double x, double y) const = 0;
  virtual ShapeDesc describe() const { return {}; }
};
std::unique_ptr<Shape> make_polygon(std::vector<std::array<double, 2>> pts);
std::vector<std::array<double, 2>> section_outline(int corner, double b, double cx, double cy, double d);
// Bisection of the first crossing along a link, 40 halvings, clamped to [1e-9, 1]
// (src/geometry.rs:268-279).
double cut_fraction(const Shape& s, double x, double y, double dx, double dy);

// ---- builder -------------------------------------------------------------------------------
using VelocityFn = std::function<std::array<double, 2>(double, double)>;
VelocityFn velocity_from_desc(const vl_velocity* v);

enum class CellKind : uint8_t { Fluid, Wall, Pressure, Radiating };
struct KindDesc {
  CellKind tag = CellKind::Fluid;
  VelocityFn velocity;  // walls: null = at rest
  int vgroup = 0;
  int fgroup = -1;
  std::unique_ptr<Shape> outline;
  double rho = 1.0;  // open boundaries
};

struct Domain {  // the reference's Builder fields (src/sim.rs:70-81)
  int64_t nx = 0, ny = 0;
  bool periodic_x = false, periodic_y = false;
  double tau = 1.0;
  int collision = VL_TRT;
  double magic = 3.0 / 16.0;
  double accel[2] = {0.0, 0.0};
  std::vector<KindDesc> kinds;  // kinds[0] is the fluid
  std::vector<uint8_t> cell;    // row-major nx*ny kind ids
};

// ---- boundary tables (what the device sees) ------------------------------------------------
// Boundary nodes are fluid nodes with at least one link into a non-fluid cell.  They are
// kept in row-major creation order, each owning a contiguous run of links in link order;
// that order fixes every serial sum the reference performs (src/sim.rs:425-508).
enum : uint8_t { LINK_WALL = 0, LINK_PRESSURE = 1 };
struct Tables {
  int64_t nx = 0, ny = 0, fluid_nodes = 0;
  std::vector<uint8_t> flag;          // nx*ny: 0 interior fluid, 1 boundary fluid, 2 not fluid
  std::vector<int32_t> bnode_x, bnode_y, bnode_first;  // bnode_first has nb+1 entries
  // links (struct of arrays)
  std::vector<uint8_t> l_dir, l_kind, l_vgroup, l_flags;  // l_flags bit0 radiating, bit1 has behind
  std::vector<int8_t> l_nx, l_ny;
  std::vector<int32_t> l_fslot;       // index into the force-contribution list, -1 none
  std::vector<double> l_q, l_uwx, l_uwy, l_cu6, l_rho;
  // pressure-boundary neighbours for the outlet controller (src/sim.rs:298-300)
  std::vector<int32_t> p_x, p_y;
  std::vector<int8_t> p_nx, p_ny;
  // force bookkeeping: links with a force group, in link order; grouped per force group
  int force_groups = 0;
  std::vector<int32_t> f_link;        // link id of each force slot
  std::vector<int32_t> f_group_first; // force_groups+1 offsets into the slot list
};

vl_status build_tables(const Domain& d, Tables& t);

// Omega pair for a collision model (src/lattice.rs:39-44).
void omegas(int collision, double magic, double tau, double& op, double& om);

// Smooth inflow ramp (src/cases/mod.rs:11-18).
double ramp(uint64_t t, uint64_t steps);

// Boundary nodes whose links use the outlet controller first (in creation order, so boundary
// node k < npb is outlet-controller node k); the rest after.  Link order inside a node and
// force slots are unchanged.  Returns npb.
int32_t outlet_nodes_first(Tables& t);
bool outlet_nodes_match(const Tables& t, int32_t npb);

// Channel::new's domain (src/cases/channel.rs:89-125); `body` is consumed.
vl_builder* channel_builder(const vl_channel_spec* sp, vl_shape* body, vl_velocity* inlet_out);
// The body kick of step t (src/cases/channel.rs:140-146).
double channel_kick(const vl_kick& k, bool has_kick, uint64_t t);

}  // namespace vl

struct vl_shape {
  std::unique_ptr<vl::Shape> s;
};
struct vl_builder {
  vl::Domain d;
};
