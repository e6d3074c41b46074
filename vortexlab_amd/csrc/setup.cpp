// Host set-up: shapes, builder and boundary tables.  See setup.h.
#include "setup.h"

#include <algorithm>
#include <cmath>
#include <cstring>

#include "lbm_math.h"

namespace vl {

static thread_local std::string t_error;
vl_status set_error(vl_status st, const std::string& msg) {
  t_error = msg;
  return st;
}
const char* last_error() { return t_error.c_str(); }

// ------------------------------------------------------------------------------------------
// Shapes (src/geometry.rs).  The predicates are the reference's, term for term, because the
// cut fractions q feed the wall rule and must come out identical.
// ------------------------------------------------------------------------------------------
namespace {
struct Disc final : Shape {  // geometry.rs:16-27
  double cx, cy, r;
  Disc(double a, double b, double c) : cx(a), cy(b), r(c) {}
  bool inside(double x, double y) const override {
    double dx = x - cx, dy = y - cy;
    return dx * dx + dy * dy < r * r;
  }
  ShapeDesc describe() const override {
    ShapeDesc d;
    d.kind = ShapeDesc::CIRCLE;
    d.cx = cx;
    d.cy = cy;
    d.r = r;
    return d;
  }
};
struct EvenOdd final : Shape {  // geometry.rs:30-49
  std::vector<std::array<double, 2>> p;
  explicit EvenOdd(std::vector<std::array<double, 2>> v) : p(std::move(v)) {}
  bool inside(double x, double y) const override {
    bool in = false;
    for (size_t i = 0, j = p.size() - 1; i < p.size(); j = i++) {
      const double xi = p[i][0], yi = p[i][1], xj = p[j][0], yj = p[j][1];
      if ((yi > y) != (yj > y) && x < (xj - xi) * (y - yi) / (yj - yi) + xi) in = !in;
    }
    return in;
  }
  ShapeDesc describe() const override {
    ShapeDesc d;
    d.kind = ShapeDesc::POLYGON;
    d.pts = p;
    return d;
  }
};
struct RoundSq final : Shape {  // geometry.rs:52-66
  double cx, cy, half, r;
  RoundSq(double a, double b, double h, double rr) : cx(a), cy(b), half(h), r(rr) {}
  bool inside(double x, double y) const override {
    const double ax = std::fabs(x - cx), ay = std::fabs(y - cy), core = half - r;
    const double dx = std::max(ax - core, 0.0), dy = std::max(ay - core, 0.0);
    return ax < half && ay < half && dx * dx + dy * dy < r * r;
  }
};
struct Turned final : Shape {  // geometry.rs:69-82
  std::unique_ptr<Shape> in;
  double cx, cy, s, c;
  Turned(std::unique_ptr<Shape> i, double a, double b, double ang)
      : in(std::move(i)), cx(a), cy(b), s(std::sin(ang)), c(std::cos(ang)) {}
  bool inside(double x, double y) const override {
    const double dx = x - cx, dy = y - cy;
    return in->inside(cx + c * dx + s * dy, cy - s * dx + c * dy);
  }
};
struct Moved final : Shape {  // geometry.rs:85-95
  std::unique_ptr<Shape> in;
  double dx, dy;
  Moved(std::unique_ptr<Shape> i, double a, double b) : in(std::move(i)), dx(a), dy(b) {}
  bool inside(double x, double y) const override { return in->inside(x - dx, y - dy); }
};
struct Band final : Shape {  // src/cases/poiseuille.rs:12-21
  double y0, y1;
  Band(double a, double b) : y0(a), y1(b) {}
  bool inside(double, double y) const override { return y <= y0 || y >= y1; }
};
struct UserShape final : Shape {
  vl_inside_fn fn;
  void* ctx;
  UserShape(vl_inside_fn f, void* c) : fn(f), ctx(c) {}
  bool inside(double x, double y) const override { return fn(ctx, x, y) != 0; }
};
}  // namespace

std::unique_ptr<Shape> make_polygon(std::vector<std::array<double, 2>> pts) {
  return std::make_unique<EvenOdd>(std::move(pts));
}

// Vertices of a square section with modified corners, counter-clockwise: the top-right
// corner's outline (offsets in units of the side) rotated by 0, 90, 180, 270 degrees
// (src/geometry.rs:215-248).
std::vector<std::array<double, 2>> section_outline(int corner, double b, double cx, double cy, double d) {
  std::vector<std::array<double, 2>> corner_pts;
  if (corner == 1)
    corner_pts = {{0.0, -b}, {-b, 0.0}};
  else if (corner == 3)
    corner_pts = {{0.0, -b}, {-b, -b}, {-b, 0.0}};
  else if (corner == 4)
    corner_pts = {{0.0, -2.0 * b}, {-b, -2.0 * b}, {-b, -b}, {-2.0 * b, -b}, {-2.0 * b, 0.0}};
  else
    corner_pts = {{0.0, 0.0}};
  const double h = 0.5 * d;
  std::vector<std::array<double, 2>> out;
  out.reserve(4 * corner_pts.size());
  for (int quarter = 0; quarter < 4; ++quarter)
    for (const auto& p : corner_pts) {
      const double x = h + p[0] * d, y = h + p[1] * d;
      const double rx = quarter == 0 ? x : quarter == 1 ? -y : quarter == 2 ? -x : y;
      const double ry = quarter == 0 ? y : quarter == 1 ? x : quarter == 2 ? -y : -x;
      out.push_back({cx + rx, cy + ry});
    }
  return out;
}

double cut_fraction(const Shape& s, double x, double y, double dx, double dy) {
  double lo = 0.0, hi = 1.0;
  for (int k = 0; k < 40; ++k) {
    const double mid = 0.5 * (lo + hi);
    (s.inside(x + mid * dx, y + mid * dy) ? hi : lo) = mid;
  }
  return std::min(std::max(0.5 * (lo + hi), 1e-9), 1.0);
}

// ------------------------------------------------------------------------------------------
// Velocity descriptors -> closures (the reference's VelocityFn, src/sim.rs:33-34).
// ------------------------------------------------------------------------------------------
VelocityFn velocity_from_desc(const vl_velocity* v) {
  if (!v) return nullptr;
  const vl_velocity c = *v;
  switch (c.kind) {
    case VL_VEL_CONST:
      return [c](double, double) { return std::array<double, 2>{c.p[0], c.p[1]}; };
    case VL_VEL_PARABOLA:  // channel profile (src/cases/channel.rs:95-103)
      return [c](double, double y) {
        const double u_mean = c.p[0], h = c.p[1];
        const double s = std::min(std::max(y - 0.5, 0.0), h);
        return std::array<double, 2>{6.0 * u_mean * s * (h - s) / (h * h), 0.0};
      };
    case VL_VEL_SPIN:  // body twist (src/cases/channel.rs:122)
      return [c](double x, double y) { return std::array<double, 2>{-(y - c.p[1]), x - c.p[0]}; };
    case VL_VEL_CALLBACK:
      if (!c.fn) return nullptr;
      return [c](double x, double y) {
        std::array<double, 2> o{0.0, 0.0};
        c.fn(c.ctx, x, y, o.data());
        return o;
      };
    default:
      return nullptr;
  }
}

void omegas(int collision, double magic, double tau, double& op, double& om) {
  op = 1.0 / tau;
  om = collision == VL_BGK ? 1.0 / tau : 1.0 / (0.5 + magic / (tau - 0.5));
}

double ramp(uint64_t t, uint64_t steps) {
  if (steps == 0 || t >= steps) return 1.0;
  const double s = (double)t / (double)steps;
  return s * s * (3.0 - 2.0 * s);
}

// ------------------------------------------------------------------------------------------
// Boundary tables (src/sim.rs:219-328).  One row-major scan: a fluid node whose eight
// neighbours are all fluid is interior (flag 0) and is handled by the bulk kernel; any other
// fluid node becomes a boundary node owning one link per neighbour that is not fluid.
// ------------------------------------------------------------------------------------------
vl_status build_tables(const Domain& d, Tables& t) {
  const int64_t nx = d.nx, ny = d.ny;
  if (!(d.tau > 0.5)) return set_error(VL_EINVAL, "tau must exceed 1/2 (viscosity must be positive)");
  if (nx * ny >= (int64_t)1 << 31) return set_error(VL_EINVAL, "grid too large (nx*ny must be < 2^31)");
  t = Tables{};
  t.nx = nx;
  t.ny = ny;
  t.flag.assign((size_t)(nx * ny), 2);
  // Neighbour lookup with optional periodic wrap; false when it leaves a closed edge.
  auto neighbour = [&](int64_t x, int64_t y, int64_t& idx) {
    if (x < 0 || x >= nx) {
      if (!d.periodic_x) return false;
      x = (x % nx + nx) % nx;
    }
    if (y < 0 || y >= ny) {
      if (!d.periodic_y) return false;
      y = (y % ny + ny) % ny;
    }
    idx = y * nx + x;
    return true;
  };
  auto is_open = [&](uint8_t k) {
    return d.kinds[k].tag == CellKind::Pressure || d.kinds[k].tag == CellKind::Radiating;
  };
  std::vector<std::vector<int32_t>> group_links;  // per force group, link ids in link order
  for (int64_t y = 0; y < ny; ++y)
    for (int64_t x = 0; x < nx; ++x) {
      const int64_t idx = y * nx + x;
      if (d.cell[idx] != 0) continue;
      ++t.fluid_nodes;
      const size_t first = t.l_dir.size();
      for (int j = 1; j < Q; ++j) {
        int64_t to;
        if (!neighbour(x + cx(j), y + cy(j), to))
          return set_error(VL_EINVAL, "fluid cell (" + std::to_string(x) + "," + std::to_string(y) +
                                          ") touches a non-periodic edge; close the domain with wall or pressure cells");
        const uint8_t kid = d.cell[to];
        if (kid == 0) continue;
        const KindDesc& k = d.kinds[kid];
        int64_t back_idx;
        const bool behind = neighbour(x - cx(j), y - cy(j), back_idx) && d.cell[back_idx] == 0;
        uint8_t kind = LINK_WALL, vgroup = 0, flags = behind ? 2 : 0;
        int8_t nxo = 0, nyo = 0;
        double q = 0.5, uwx = 0.0, uwy = 0.0, cu6 = 0.0, rho = 1.0;
        int fgroup = -1;
        if (k.tag == CellKind::Pressure || k.tag == CellKind::Radiating) {
          kind = LINK_PRESSURE;
          rho = k.rho;
          const bool radiating = k.tag == CellKind::Radiating;
          if (radiating) flags |= 1;
          // Outward normal along x if the cell beside the node in x is open too, else along y.
          int64_t side;
          if (neighbour(x + cx(j), y, side) && is_open(d.cell[side]))
            nxo = (int8_t)cx(j);
          else
            nyo = (int8_t)cy(j);
          if (!radiating && (t.p_x.empty() || t.p_x.back() != x || t.p_y.back() != y)) {
            t.p_x.push_back((int32_t)x);
            t.p_y.push_back((int32_t)y);
            t.p_nx.push_back(nxo);
            t.p_ny.push_back(nyo);
          }
        } else {
          if (k.vgroup < 0 || k.vgroup >= 4) return set_error(VL_EINVAL, "wall vgroup must be 0..3");
          vgroup = (uint8_t)k.vgroup;
          if (k.outline) q = cut_fraction(*k.outline, (double)x, (double)y, (double)cx(j), (double)cy(j));
          if (k.velocity) {
            const auto uw = k.velocity((double)x + q * (double)cx(j), (double)y + q * (double)cy(j));
            uwx = uw[0];
            uwy = uw[1];
            cu6 = 6.0 * wgt(j) * ((double)cx(j) * uwx + (double)cy(j) * uwy);
          }
          if (k.fgroup >= 0) {
            fgroup = k.fgroup;
            if ((int)group_links.size() <= fgroup) group_links.resize(fgroup + 1);
            group_links[fgroup].push_back((int32_t)t.l_dir.size());
          }
        }
        t.l_dir.push_back((uint8_t)j);
        t.l_kind.push_back(kind);
        t.l_vgroup.push_back(vgroup);
        t.l_flags.push_back(flags);
        t.l_nx.push_back(nxo);
        t.l_ny.push_back(nyo);
        t.l_fslot.push_back(fgroup);  // rewritten to a slot index below
        t.l_q.push_back(q);
        t.l_uwx.push_back(uwx);
        t.l_uwy.push_back(uwy);
        t.l_cu6.push_back(cu6);
        t.l_rho.push_back(rho);
      }
      if (t.l_dir.size() > first) {
        t.flag[idx] = 1;
        t.bnode_x.push_back((int32_t)x);
        t.bnode_y.push_back((int32_t)y);
        t.bnode_first.push_back((int32_t)first);
      } else {
        t.flag[idx] = 0;
      }
    }
  t.bnode_first.push_back((int32_t)t.l_dir.size());
  t.force_groups = (int)group_links.size();
  t.f_group_first.assign(1, 0);
  for (auto& g : group_links) {
    for (int32_t l : g) {
      t.l_fslot[l] = (int32_t)t.f_link.size();
      t.f_link.push_back(l);
    }
    t.f_group_first.push_back((int32_t)t.f_link.size());
  }
  return VL_OK;
}

int32_t outlet_nodes_first(Tables& tb) {
  const size_t nb = tb.bnode_x.size();
  std::vector<char> needs(nb, 0);
  for (size_t b = 0; b < nb; ++b)
    for (int32_t l = tb.bnode_first[b]; l < tb.bnode_first[b + 1]; ++l)
      if (tb.l_kind[l] == LINK_PRESSURE && !(tb.l_flags[l] & 1)) needs[b] = 1;
  std::vector<int32_t> order;
  order.reserve(nb);
  for (size_t b = 0; b < nb; ++b)
    if (needs[b]) order.push_back((int32_t)b);
  const int32_t npb = (int32_t)order.size();
  for (size_t b = 0; b < nb; ++b)
    if (!needs[b]) order.push_back((int32_t)b);
  Tables r = tb;
  r.bnode_x.clear();
  r.bnode_y.clear();
  r.bnode_first.assign(1, 0);
  auto clear_all = [](auto&... v) { (v.clear(), ...); };
  clear_all(r.l_dir, r.l_kind, r.l_vgroup, r.l_flags, r.l_nx, r.l_ny, r.l_fslot, r.l_q, r.l_uwx, r.l_uwy, r.l_cu6,
            r.l_rho);
  for (int32_t b : order) {
    r.bnode_x.push_back(tb.bnode_x[b]);
    r.bnode_y.push_back(tb.bnode_y[b]);
    for (int32_t l = tb.bnode_first[b]; l < tb.bnode_first[b + 1]; ++l) {
      r.l_dir.push_back(tb.l_dir[l]);
      r.l_kind.push_back(tb.l_kind[l]);
      r.l_vgroup.push_back(tb.l_vgroup[l]);
      r.l_flags.push_back(tb.l_flags[l]);
      r.l_nx.push_back(tb.l_nx[l]);
      r.l_ny.push_back(tb.l_ny[l]);
      r.l_fslot.push_back(tb.l_fslot[l]);
      r.l_q.push_back(tb.l_q[l]);
      r.l_uwx.push_back(tb.l_uwx[l]);
      r.l_uwy.push_back(tb.l_uwy[l]);
      r.l_cu6.push_back(tb.l_cu6[l]);
      r.l_rho.push_back(tb.l_rho[l]);
    }
    r.bnode_first.push_back((int32_t)r.l_dir.size());
  }
  tb = std::move(r);
  return npb;
}

bool outlet_nodes_match(const Tables& tb, int32_t npb) {
  if ((size_t)npb != tb.p_x.size()) return false;
  for (int32_t k = 0; k < npb; ++k)
    if (tb.bnode_x[k] != tb.p_x[k] || tb.bnode_y[k] != tb.p_y[k]) return false;
  return true;
}

}  // namespace vl

// ==========================================================================================
// C ABI: shapes and builder
// ==========================================================================================
using namespace vl;

extern "C" {

const char* vl_last_error(void) { return last_error(); }
int vl_abi_version(void) { return VL_ABI_VERSION; }

static vl_shape* box(std::unique_ptr<Shape> s) {
  auto* o = new vl_shape;
  o->s = std::move(s);
  return o;
}
static std::unique_ptr<Shape> unbox(vl_shape* s) {
  std::unique_ptr<Shape> r = std::move(s->s);
  delete s;
  return r;
}

vl_shape* vl_shape_circle(double cx, double cy, double r) { return box(std::make_unique<Disc>(cx, cy, r)); }
vl_shape* vl_shape_polygon(const double* xy, size_t n) {
  if (!xy || n < 3) {
    set_error(VL_EINVAL, "polygon needs at least 3 points");
    return nullptr;
  }
  std::vector<std::array<double, 2>> p(n);
  for (size_t i = 0; i < n; ++i) p[i] = {xy[2 * i], xy[2 * i + 1]};
  return box(make_polygon(std::move(p)));
}
vl_shape* vl_shape_rounded_square(double cx, double cy, double half, double r) {
  return box(std::make_unique<RoundSq>(cx, cy, half, r));
}
vl_shape* vl_shape_rotated(vl_shape* inner, double cx, double cy, double angle) {
  if (!inner) return nullptr;
  return box(std::make_unique<Turned>(unbox(inner), cx, cy, angle));
}
vl_shape* vl_shape_translated(vl_shape* inner, double dx, double dy) {
  if (!inner) return nullptr;
  return box(std::make_unique<Moved>(unbox(inner), dx, dy));
}
vl_shape* vl_shape_section(int corner, double b, double cx, double cy, double d, double angle) {
  if (corner < 0 || corner > 4) {
    set_error(VL_EINVAL, "corner must be 0..4");
    return nullptr;
  }
  std::unique_ptr<Shape> base;
  if (corner == 2)
    base = std::make_unique<RoundSq>(cx, cy, 0.5 * d, b * d);
  else
    base = make_polygon(section_outline(corner, b, cx, cy, d));
  if (angle != 0.0) base = std::make_unique<Turned>(std::move(base), cx, cy, angle);
  return box(std::move(base));
}
vl_shape* vl_shape_outside_band(double y0, double y1) { return box(std::make_unique<Band>(y0, y1)); }
vl_shape* vl_shape_callback(vl_inside_fn fn, void* ctx) {
  if (!fn) return nullptr;
  return box(std::make_unique<UserShape>(fn, ctx));
}
int vl_shape_inside(const vl_shape* s, double x, double y) { return s && s->s->inside(x, y) ? 1 : 0; }
double vl_cut_fraction(const vl_shape* s, double x, double y, double dx, double dy) {
  return s ? cut_fraction(*s->s, x, y, dx, dy) : NAN;
}
void vl_shape_free(vl_shape* s) { delete s; }

vl_builder* vl_builder_new(size_t nx, size_t ny) {
  if (nx < 3 || ny < 3) {
    set_error(VL_EINVAL, "grid must be at least 3 x 3");
    return nullptr;
  }
  auto* b = new vl_builder;
  b->d.nx = (int64_t)nx;
  b->d.ny = (int64_t)ny;
  b->d.kinds.emplace_back();
  b->d.cell.assign(nx * ny, 0);
  return b;
}
vl_status vl_builder_config(vl_builder* b, int px, int py, double tau, int collision, double magic, double ax,
                            double ay) {
  if (!b) return set_error(VL_EINVAL, "null builder");
  if (collision != VL_BGK && collision != VL_TRT) return set_error(VL_EINVAL, "unknown collision model");
  b->d.periodic_x = px != 0;
  b->d.periodic_y = py != 0;
  b->d.tau = tau;
  b->d.collision = collision;
  b->d.magic = magic;
  b->d.accel[0] = ax;
  b->d.accel[1] = ay;
  return VL_OK;
}
static int push_kind(vl_builder* b, KindDesc&& k) {
  if (b->d.kinds.size() >= 255) {
    set_error(VL_EINVAL, "at most 254 cell kinds");
    return -1;
  }
  b->d.kinds.push_back(std::move(k));
  return (int)b->d.kinds.size() - 1;
}
int vl_builder_add_wall(vl_builder* b, const vl_velocity* vel, int vgroup, int fgroup, vl_shape* shape) {
  if (!b) return -1;
  if (vgroup < 0 || vgroup >= 4) {
    if (shape) vl_shape_free(shape);
    set_error(VL_EINVAL, "wall vgroup must be 0..3");
    return -1;
  }
  KindDesc k;
  k.tag = CellKind::Wall;
  k.velocity = (vel && vel->kind != VL_VEL_REST) ? velocity_from_desc(vel) : nullptr;
  k.vgroup = vgroup;
  k.fgroup = fgroup < 0 ? -1 : fgroup;
  if (shape) k.outline = unbox(shape);
  return push_kind(b, std::move(k));
}
int vl_builder_add_pressure(vl_builder* b, double rho) {
  if (!b) return -1;
  KindDesc k;
  k.tag = CellKind::Pressure;
  k.rho = rho;
  return push_kind(b, std::move(k));
}
int vl_builder_add_radiating(vl_builder* b, double rho) {
  if (!b) return -1;
  KindDesc k;
  k.tag = CellKind::Radiating;
  k.rho = rho;
  return push_kind(b, std::move(k));
}
vl_status vl_builder_fill_rect(vl_builder* b, size_t x0, size_t x1, size_t y0, size_t y1, int kind) {
  if (!b) return set_error(VL_EINVAL, "null builder");
  if (kind < 0 || (size_t)kind >= b->d.kinds.size()) return set_error(VL_EINVAL, "unknown kind id");
  if (x1 >= (size_t)b->d.nx || y1 >= (size_t)b->d.ny) return set_error(VL_EINVAL, "rectangle outside the grid");
  for (size_t y = y0; y <= y1; ++y)
    std::memset(&b->d.cell[y * b->d.nx + x0], kind, x1 >= x0 ? x1 - x0 + 1 : 0);
  return VL_OK;
}
vl_status vl_builder_fill_shape(vl_builder* b, int kind) {
  if (!b) return set_error(VL_EINVAL, "null builder");
  if (kind <= 0 || (size_t)kind >= b->d.kinds.size() || b->d.kinds[kind].tag != CellKind::Wall ||
      !b->d.kinds[kind].outline)
    return set_error(VL_EINVAL, "fill_shape needs a wall kind with an outline");
  const Shape& s = *b->d.kinds[kind].outline;
  for (int64_t y = 0; y < b->d.ny; ++y)
    for (int64_t x = 0; x < b->d.nx; ++x)
      if (s.inside((double)x, (double)y)) b->d.cell[y * b->d.nx + x] = (uint8_t)kind;
  return VL_OK;
}
void vl_builder_free(vl_builder* b) { delete b; }

}  // extern "C"
