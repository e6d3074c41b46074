// CPU oracle for the D2Q9 time step: a restatement of the reference crate's solver in C++.
//
// TEST INFRASTRUCTURE ONLY -- the checker, never the product.  Only tests/, the smoke()
// check and bench.py's cpu_baseline leg load it.
//
// Every function follows the reference's floating-point operation order exactly (left to
// right association, the same divisions, no FMA contraction: build with -ffp-contract=off
// and without fast-math).  Rust never contracts a*b+c, so with these flags the arithmetic
// is the same IEEE-754 sequence and the published FNV-1a state hashes are reproduced bit
// for bit.  Each function names the reference lines it restates.
//
// Pinned by: the published state hashes 6b6814b976bbbaff (n=40, 3000 steps) and
// c93e26bfa33474ee (n=80, 1000 steps) with fluid-node counts 143056 / 572256
// (docs/report/index.html:59); the published Poiseuille L2 table (docs/report/index.html,
// section 1); the reference's tolerance tests (tests/solver.rs, src/lattice.rs tests,
// src/signal.rs tests) re-expressed under tests/.

#include "oracle.h"

#include <algorithm>
#include <array>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstdint>
#include <cstring>
#include <functional>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

namespace {

thread_local std::string g_err;
vl_status fail(vl_status st, const std::string& msg) {
  g_err = msg;
  return st;
}

// ---- src/lattice.rs:16-24 -------------------------------------------------------------
constexpr int Q = 9;
constexpr int CX[Q] = {0, 1, 0, -1, 0, 1, -1, -1, 1};
constexpr int CY[Q] = {0, 0, 1, 0, -1, 1, 1, -1, -1};
const double W[Q] = {4.0 / 9.0, 1.0 / 9.0, 1.0 / 9.0, 1.0 / 9.0, 1.0 / 9.0,
                     1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0};
constexpr int OPP[Q] = {0, 3, 4, 1, 2, 7, 8, 5, 6};
const double CS2 = 1.0 / 3.0;
constexpr int PAIRS[4][2] = {{1, 3}, {2, 4}, {5, 7}, {6, 8}};

using Pop = std::array<double, Q>;

// src/lattice.rs:39-44
void omegas(int collision, double magic, double tau, double& op, double& om) {
  if (collision == VL_BGK) {
    op = 1.0 / tau;
    om = 1.0 / tau;
  } else {
    op = 1.0 / tau;
    om = 1.0 / (0.5 + magic / (tau - 0.5));
  }
}

// src/lattice.rs:59-64
inline void moments(const Pop& f, double& rho, double& jx, double& jy) {
  rho = f[0] + f[1] + f[2] + f[3] + f[4] + f[5] + f[6] + f[7] + f[8];
  jx = f[1] - f[3] + f[5] - f[6] - f[7] + f[8];
  jy = f[2] - f[4] + f[5] + f[6] - f[7] - f[8];
}

// src/lattice.rs:68-76
inline Pop equilibrium(double rho, double ux, double uy) {
  double usq = 1.5 * (ux * ux + uy * uy);
  Pop out{};
  for (int i = 0; i < Q; ++i) {
    double cu = 3.0 * ((double)CX[i] * ux + (double)CY[i] * uy);
    out[i] = W[i] * rho * (1.0 + cu + 0.5 * cu * cu - usq);
  }
  return out;
}

// src/lattice.rs:80-95
inline Pop collide_bgk(const Pop& f, double omega) {
  double rho, jx, jy;
  moments(f, rho, jx, jy);
  double inv = 1.0 / rho;
  Pop feq = equilibrium(rho, jx * inv, jy * inv);
  Pop out{};
  double moving = 0.0;
  for (int i = 1; i < Q; ++i) {
    out[i] = f[i] - omega * (f[i] - feq[i]);
    moving += feq[i];
  }
  out[0] = f[0] - omega * (f[0] - (rho - moving));
  return out;
}

// src/lattice.rs:100-119
inline Pop collide_trt(const Pop& f, double op, double om) {
  double rho, jx, jy;
  moments(f, rho, jx, jy);
  double inv = 1.0 / rho;
  double ux = jx * inv, uy = jy * inv;
  double usq = 1.5 * (ux * ux + uy * uy);
  Pop out{};
  double moving = 0.0;
  for (int p = 0; p < 4; ++p) {
    int a = PAIRS[p][0], b = PAIRS[p][1];
    double cu = 3.0 * ((double)CX[a] * ux + (double)CY[a] * uy);
    double wr = W[a] * rho;
    double eq = wr * (1.0 + 0.5 * cu * cu - usq);
    double dp = op * (0.5 * (f[a] + f[b]) - eq);
    double dm = om * (0.5 * (f[a] - f[b]) - wr * cu);
    out[a] = f[a] - dp - dm;
    out[b] = f[b] - dp + dm;
    moving += 2.0 * eq;
  }
  out[0] = f[0] - op * (f[0] - (rho - moving));
  return out;
}

// src/lattice.rs:127-156
inline Pop collide_forced(const Pop& f, double op, double om, double gx, double gy) {
  double rho, jx, jy;
  moments(f, rho, jx, jy);
  double inv = 1.0 / rho;
  double ux = jx * inv + 0.5 * gx, uy = jy * inv + 0.5 * gy;
  double fx = rho * gx, fy = rho * gy;
  double usq = 1.5 * (ux * ux + uy * uy);
  double uf = 3.0 * (ux * fx + uy * fy);
  double kp = 1.0 - 0.5 * op, km = 1.0 - 0.5 * om;
  Pop out{};
  double moving = 0.0, source = 0.0;
  for (int p = 0; p < 4; ++p) {
    int a = PAIRS[p][0], b = PAIRS[p][1];
    double cx = (double)CX[a], cy = (double)CY[a];
    double cu = 3.0 * (cx * ux + cy * uy);
    double cf = 3.0 * (cx * fx + cy * fy);
    double wr = W[a] * rho;
    double eq = wr * (1.0 + 0.5 * cu * cu - usq);
    double dp = op * (0.5 * (f[a] + f[b]) - eq);
    double dm = om * (0.5 * (f[a] - f[b]) - wr * cu);
    double sp = kp * W[a] * (cu * cf - uf);
    double sm = km * W[a] * cf;
    out[a] = f[a] - dp - dm + sp + sm;
    out[b] = f[b] - dp + dm + sp - sm;
    moving += 2.0 * eq;
    source += 2.0 * sp;
  }
  out[0] = f[0] - op * (f[0] - (rho - moving)) - source;
  return out;
}

// ---- src/geometry.rs --------------------------------------------------------------------
struct Shape {
  virtual ~Shape() = default;
  virtual bool inside(double x, double y) const = 0;
};

struct Circle : Shape {  // geometry.rs:16-27
  double cx, cy, r;
  bool inside(double x, double y) const override {
    double dx = x - cx, dy = y - cy;
    return dx * dx + dy * dy < r * r;
  }
};

struct Polygon : Shape {  // geometry.rs:30-49 (even-odd rule)
  std::vector<std::array<double, 2>> pts;
  bool inside(double x, double y) const override {
    size_t n = pts.size();
    bool in = false;
    size_t j = n - 1;
    for (size_t i = 0; i < n; ++i) {
      double xi = pts[i][0], yi = pts[i][1];
      double xj = pts[j][0], yj = pts[j][1];
      if ((yi > y) != (yj > y) && x < (xj - xi) * (y - yi) / (yj - yi) + xi) in = !in;
      j = i;
    }
    return in;
  }
};

struct RoundedSquare : Shape {  // geometry.rs:52-66
  double cx, cy, half, r;
  bool inside(double x, double y) const override {
    double core = half - r;
    double dx = std::max(std::fabs(x - cx) - core, 0.0);
    double dy = std::max(std::fabs(y - cy) - core, 0.0);
    return std::fabs(x - cx) < half && std::fabs(y - cy) < half && dx * dx + dy * dy < r * r;
  }
};

struct Rotated : Shape {  // geometry.rs:69-82
  std::unique_ptr<Shape> inner;
  double cx, cy, angle;
  bool inside(double x, double y) const override {
    double s = std::sin(angle), c = std::cos(angle);
    double dx = x - cx, dy = y - cy;
    return inner->inside(cx + c * dx + s * dy, cy - s * dx + c * dy);
  }
};

struct Translated : Shape {  // geometry.rs:85-95
  std::unique_ptr<Shape> inner;
  double dx, dy;
  bool inside(double x, double y) const override { return inner->inside(x - dx, y - dy); }
};

struct OutsideBand : Shape {  // src/cases/poiseuille.rs:12-21
  double y0, y1;
  bool inside(double, double y) const override { return y <= y0 || y >= y1; }
};

struct Callback : Shape {
  vl_inside_fn fn;
  void* ctx;
  bool inside(double x, double y) const override { return fn(ctx, x, y) != 0; }
};

// geometry.rs:217-226 (corner profile) and :230-248 (section polygon)
std::vector<std::array<double, 2>> section_polygon(int corner, double b, double cx, double cy,
                                                   double d) {
  std::vector<std::array<double, 2>> prof;
  switch (corner) {
    case 1: prof = {{0.0, -b}, {-b, 0.0}}; break;
    case 3: prof = {{0.0, -b}, {-b, -b}, {-b, 0.0}}; break;
    case 4:
      prof = {{0.0, -2.0 * b}, {-b, -2.0 * b}, {-b, -b}, {-2.0 * b, -b}, {-2.0 * b, 0.0}};
      break;
    default: prof = {{0.0, 0.0}}; break;
  }
  double h = 0.5 * d;
  std::vector<std::array<double, 2>> pts;
  for (int k = 0; k < 4; ++k) {
    for (auto& p : prof) {
      double x = h + p[0] * d, y = h + p[1] * d;
      double rx, ry;
      switch (k) {
        case 0: rx = x; ry = y; break;
        case 1: rx = -y; ry = x; break;
        case 2: rx = -x; ry = -y; break;
        default: rx = y; ry = -x; break;
      }
      pts.push_back({cx + rx, cy + ry});
    }
  }
  return pts;
}

// geometry.rs:268-279
double cut_fraction(const Shape& s, double x, double y, double dx, double dy) {
  double lo = 0.0, hi = 1.0;
  for (int it = 0; it < 40; ++it) {
    double mid = 0.5 * (lo + hi);
    if (s.inside(x + mid * dx, y + mid * dy))
      hi = mid;
    else
      lo = mid;
  }
  double v = 0.5 * (lo + hi);
  if (v < 1e-9) v = 1e-9;
  if (v > 1.0) v = 1.0;
  return v;
}

// ---- src/sim.rs ---------------------------------------------------------------------------
using VelFn = std::function<std::array<double, 2>(double, double)>;

// vl_velocity -> closure, matching the reference closures it stands for.
VelFn make_velocity(const vl_velocity* v) {
  if (!v || v->kind == VL_VEL_REST) return nullptr;
  vl_velocity c = *v;
  switch (c.kind) {
    case VL_VEL_CONST:
      return [c](double, double) { return std::array<double, 2>{c.p[0], c.p[1]}; };
    case VL_VEL_PARABOLA:  // channel.rs:95-103, inlet closure :104
      return [c](double, double y) {
        double u_mean = c.p[0], h = c.p[1];
        double s = y - 0.5;
        if (s < 0.0) s = 0.0;
        if (s > h) s = h;
        return std::array<double, 2>{6.0 * u_mean * s * (h - s) / (h * h), 0.0};
      };
    case VL_VEL_SPIN:  // channel.rs:122
      return [c](double x, double y) {
        return std::array<double, 2>{-(y - c.p[1]), x - c.p[0]};
      };
    case VL_VEL_CALLBACK:
      return [c](double x, double y) {
        std::array<double, 2> o{0.0, 0.0};
        c.fn(c.ctx, x, y, o.data());
        return o;
      };
  }
  return nullptr;
}

enum KindTag { K_FLUID, K_WALL, K_PRESSURE, K_RADIATING };
struct Kind {  // sim.rs:36-63
  KindTag tag = K_FLUID;
  VelFn velocity;
  int vgroup = 0;
  long fgroup = -1;
  std::unique_ptr<Shape> shape;
  double rho = 1.0;
};

constexpr size_t NONE = SIZE_MAX;
constexpr int VGROUPS = 4;

struct Link {  // sim.rs:143-163
  uint8_t dir, kind, vgroup;
  double q;
  size_t behind, fgroup;
  double uw[2];
  double cu6, rho;
  int8_t normal[2];
  bool radiating;
};
struct BNode {  // sim.rs:165-170
  size_t idx;
  size_t src[Q];
  size_t l0, l1;
};

}  // namespace

struct or_shape {
  std::unique_ptr<Shape> s;
};

struct or_builder {  // sim.rs:70-81
  size_t nx, ny;
  bool periodic_x = false, periodic_y = false;
  double tau = 1.0;
  int collision = VL_TRT;
  double magic = 3.0 / 16.0;
  double accel[2] = {0.0, 0.0};
  std::vector<Kind> kinds;
  std::vector<uint8_t> cell;
};

namespace {
// A persistent worker pool with a static split (src/pool.rs behaviour: run(f) calls f(t)
// for every thread t and blocks until all return).
class Pool {
 public:
  explicit Pool(int n) : n_(std::max(1, n)) {
    for (int t = 1; t < n_; ++t) workers_.emplace_back([this, t] { worker(t); });
  }
  ~Pool() {
    {
      std::lock_guard<std::mutex> g(m_);
      stop_ = true;
      ++gen_;
    }
    cv_.notify_all();
    for (auto& w : workers_) w.join();
  }
  int threads() const { return n_; }
  void run(const std::function<void(int)>& f) {
    if (n_ == 1) {
      f(0);
      return;
    }
    {
      std::lock_guard<std::mutex> g(m_);
      job_ = &f;
      pending_ = n_ - 1;
      ++gen_;
    }
    cv_.notify_all();
    f(0);
    std::unique_lock<std::mutex> lk(m_);
    done_.wait(lk, [this] { return pending_ == 0; });
    job_ = nullptr;
  }

 private:
  void worker(int t) {
    uint64_t seen = 0;
    for (;;) {
      const std::function<void(int)>* job;
      {
        std::unique_lock<std::mutex> lk(m_);
        cv_.wait(lk, [&] { return gen_ != seen; });
        seen = gen_;
        if (stop_) return;
        job = job_;
      }
      (*job)(t);
      {
        std::lock_guard<std::mutex> g(m_);
        if (--pending_ == 0) done_.notify_one();
      }
    }
  }
  int n_;
  std::vector<std::thread> workers_;
  std::mutex m_;
  std::condition_variable cv_, done_;
  const std::function<void(int)>* job_ = nullptr;
  int pending_ = 0;
  uint64_t gen_ = 0;
  bool stop_ = false;
};
}  // namespace

struct or_sim {  // sim.rs:180-216
  size_t nx, ny, n;
  bool periodic_x, periodic_y;
  std::vector<double> a, b;
  std::vector<uint8_t> flag;
  std::vector<BNode> bnodes;
  std::vector<Link> links;
  double op, om;
  int collision;
  double accel[2];
  std::unique_ptr<Pool> pool;
  double vscale[VGROUPS] = {1.0, 1.0, 1.0, 1.0};
  std::vector<std::array<double, 2>> force;
  bool has_target = false;
  double target = 0.0;
  double outlet_density_offset = 0.0;
  struct PNode {
    size_t idx;
    int8_t nx, ny;
  };
  std::vector<PNode> pnodes;
  uint64_t t = 0;
  size_t fluid_nodes = 0;

  Pop pops(size_t idx) const {
    Pop f;
    for (int i = 0; i < Q; ++i) f[i] = a[i * n + idx];
    return f;
  }
  // sim.rs:519-524 (and the link-pass closure :432-439, same arithmetic)
  void velocity_at(size_t idx, double& ux, double& uy) const {
    double rho, jx, jy;
    moments(pops(idx), rho, jx, jy);
    ux = jx / rho - 0.5 * accel[0];
    uy = jy / rho - 0.5 * accel[1];
  }
  Pop collide(const Pop& f) const {  // dispatch of sim.rs:407-413 / :497-503
    if (accel[0] != 0.0 || accel[1] != 0.0) return collide_forced(f, op, om, accel[0], accel[1]);
    if (collision == VL_BGK) return collide_bgk(f, op);
    return collide_trt(f, op, om);
  }
  void init(const std::function<void(size_t, size_t, double&, double&, double&)>& field);
  void step();
  void row(size_t y);
  void link_pass();
};

// sim.rs:357-371
void or_sim::init(const std::function<void(size_t, size_t, double&, double&, double&)>& field) {
  for (size_t y = 0; y < ny; ++y)
    for (size_t x = 0; x < nx; ++x) {
      size_t idx = y * nx + x;
      double rho = 1.0, ux = 0.0, uy = 0.0;
      if (flag[idx] != 2) field(x, y, rho, ux, uy);
      Pop feq = equilibrium(rho, ux, uy);
      for (int i = 0; i < Q; ++i) {
        a[i * n + idx] = feq[i];
        b[i * n + idx] = feq[i];
      }
    }
  t = 0;
}

// sim.rs:700-735: stream-and-collide for the interior fluid nodes of one row.
void or_sim::row(size_t y) {
  size_t ym = (y == 0) ? ny - 1 : y - 1;
  size_t yp = (y == ny - 1) ? 0 : y + 1;
  size_t base = y * nx, bm = ym * nx, bp = yp * nx;
  const double* A = a.data();
  double* B = b.data();
  for (size_t x = 0; x < nx; ++x) {
    if (flag[base + x] != 0) continue;
    size_t xm = (x == 0) ? nx - 1 : x - 1;
    size_t xp = (x == nx - 1) ? 0 : x + 1;
    Pop f = {A[base + x],          A[n + base + xm],     A[2 * n + bm + x],
             A[3 * n + base + xp], A[4 * n + bp + x],     A[5 * n + bm + xm],
             A[6 * n + bm + xp],   A[7 * n + bp + xp],   A[8 * n + bp + xm]};
    Pop o = collide(f);
    for (int i = 0; i < Q; ++i) B[i * n + base + x] = o[i];
  }
}

// sim.rs:391-420
void or_sim::step() {
  int nt = pool->threads();
  pool->run([&](int th) {
    for (size_t y = (size_t)th * ny / nt; y < (size_t)(th + 1) * ny / nt; ++y) row(y);
  });
  link_pass();
  std::swap(a, b);
  t += 1;
}

// sim.rs:425-508
void or_sim::link_pass() {
  for (auto& g : force) g = {0.0, 0.0};
  double drho = 0.0;
  if (has_target && !pnodes.empty()) {  // sim.rs:444-454
    double sum = 0.0;
    for (auto& p : pnodes) {
      double ux, uy;
      velocity_at(p.idx, ux, uy);
      sum += (double)p.nx * ux + (double)p.ny * uy;
    }
    drho = (sum / (double)pnodes.size() - target) / std::sqrt(CS2);
  }
  outlet_density_offset = drho;
  const double* A = a.data();
  double* B = b.data();
  for (const BNode& bn : bnodes) {
    size_t idx = bn.idx;
    Pop f;
    for (int i = 0; i < Q; ++i) f[i] = A[i * n + bn.src[i]];
    for (size_t li = bn.l0; li < bn.l1; ++li) {
      const Link& l = links[li];
      int j = l.dir;
      int i = OPP[j];
      double out = A[j * n + idx];
      double scale = vscale[l.vgroup];
      double back;
      if (l.kind == 0) {  // LINK_WALL
        double corr = l.cu6 * scale;
        if (l.q >= 0.5) {
          double k = 0.5 / l.q;
          back = k * (out - corr) + (1.0 - k) * A[i * n + idx];
        } else if (l.behind != NONE) {
          back = 2.0 * l.q * out + (1.0 - 2.0 * l.q) * A[j * n + l.behind] - corr;
        } else {
          back = out - corr;
        }
      } else {  // LINK_PRESSURE: anti-bounce-back
        double ux, uy;
        velocity_at(idx, ux, uy);
        double cu = 3.0 * ((double)CX[j] * ux + (double)CY[j] * uy);
        double rho_w = l.radiating
                           ? l.rho + ((double)l.normal[0] * ux + (double)l.normal[1] * uy) / std::sqrt(CS2)
                           : l.rho + drho;
        back = -out + 2.0 * W[j] * rho_w * (1.0 + 0.5 * cu * cu - 1.5 * (ux * ux + uy * uy));
      }
      f[i] = back;
      if (l.fgroup != NONE) {
        double ux = l.uw[0] * scale, uy = l.uw[1] * scale;
        auto& g = force[l.fgroup];
        g[0] += (double)CX[j] * (out + back) - ux * (out - back);
        g[1] += (double)CY[j] * (out + back) - uy * (out - back);
      }
    }
    Pop o = collide(f);
    for (int i = 0; i < Q; ++i) B[i * n + idx] = o[i];
  }
}

namespace {

// sim.rs:219-354
vl_status from_builder(or_builder* bld, int threads, or_sim** out) {
  size_t nx = bld->nx, ny = bld->ny, n = nx * ny;
  if (!(bld->tau > 0.5)) return fail(VL_EINVAL, "tau must exceed 1/2 (viscosity must be positive)");
  auto wrap = [&](long x, long y, size_t& res) -> bool {
    if (x < 0 || x >= (long)nx) {
      if (!bld->periodic_x) return false;
      x = ((x % (long)nx) + (long)nx) % (long)nx;
    }
    if (y < 0 || y >= (long)ny) {
      if (!bld->periodic_y) return false;
      y = ((y % (long)ny) + (long)ny) % (long)ny;
    }
    res = (size_t)y * nx + (size_t)x;
    return true;
  };
  auto s = std::make_unique<or_sim>();
  s->nx = nx;
  s->ny = ny;
  s->n = n;
  s->periodic_x = bld->periodic_x;
  s->periodic_y = bld->periodic_y;
  s->flag.assign(n, 2);
  size_t ngroups = 0;
  for (size_t y = 0; y < ny; ++y)
    for (size_t x = 0; x < nx; ++x) {
      size_t idx = y * nx + x;
      if (bld->cell[idx] != 0) continue;
      s->fluid_nodes += 1;
      size_t first = s->links.size();
      BNode bn;
      bn.idx = idx;
      for (int i = 0; i < Q; ++i) bn.src[i] = idx;
      for (int j = 1; j < Q; ++j) {
        long cx = CX[j], cy = CY[j];
        size_t target;
        if (!wrap((long)x + cx, (long)y + cy, target))
          return fail(VL_EINVAL, "fluid cell (" + std::to_string(x) + "," + std::to_string(y) +
                                     ") touches a non-periodic edge; close the domain with wall or pressure cells");
        bn.src[OPP[j]] = target;
        uint8_t kind = bld->cell[target];
        if (kind == 0) continue;
        size_t behind = NONE, bi;
        if (wrap((long)x - cx, (long)y - cy, bi) && bld->cell[bi] == 0) behind = bi;
        Link l{};
        l.dir = (uint8_t)j;
        l.kind = 0;
        l.vgroup = 0;
        l.q = 0.5;
        l.behind = behind;
        l.fgroup = NONE;
        l.uw[0] = l.uw[1] = 0.0;
        l.cu6 = 0.0;
        l.rho = 1.0;
        l.normal[0] = l.normal[1] = 0;
        l.radiating = false;
        const Kind& k = bld->kinds[kind];
        if (k.tag == K_PRESSURE || k.tag == K_RADIATING) {
          l.kind = 1;
          l.rho = k.rho;
          l.radiating = k.tag == K_RADIATING;
          size_t side;
          bool open = false;
          if (wrap((long)x + cx, (long)y, side)) {
            KindTag tg = bld->kinds[bld->cell[side]].tag;
            open = tg == K_PRESSURE || tg == K_RADIATING;
          }
          if (open) {
            l.normal[0] = (int8_t)cx;
            l.normal[1] = 0;
          } else {
            l.normal[0] = 0;
            l.normal[1] = (int8_t)cy;
          }
          if (!l.radiating && (s->pnodes.empty() || s->pnodes.back().idx != idx))
            s->pnodes.push_back({idx, l.normal[0], l.normal[1]});
        } else if (k.tag == K_WALL) {
          if (k.vgroup < 0 || k.vgroup >= VGROUPS) return fail(VL_EINVAL, "wall vgroup out of range");
          l.vgroup = (uint8_t)k.vgroup;
          if (k.shape) l.q = cut_fraction(*k.shape, (double)x, (double)y, (double)cx, (double)cy);
          if (k.velocity) {
            double wx = (double)x + l.q * (double)cx, wy = (double)y + l.q * (double)cy;
            auto uw = k.velocity(wx, wy);
            l.uw[0] = uw[0];
            l.uw[1] = uw[1];
            l.cu6 = 6.0 * W[j] * ((double)cx * l.uw[0] + (double)cy * l.uw[1]);
          }
          if (k.fgroup >= 0) {
            l.fgroup = (size_t)k.fgroup;
            ngroups = std::max(ngroups, (size_t)k.fgroup + 1);
          }
        }
        s->links.push_back(l);
      }
      if (s->links.size() > first) {
        s->flag[idx] = 1;
        bn.l0 = first;
        bn.l1 = s->links.size();
        s->bnodes.push_back(bn);
      } else {
        s->flag[idx] = 0;
      }
    }
  s->a.assign(Q * n, 0.0);
  s->b.assign(Q * n, 0.0);
  omegas(bld->collision, bld->magic, bld->tau, s->op, s->om);
  s->collision = bld->collision;
  s->accel[0] = bld->accel[0];
  s->accel[1] = bld->accel[1];
  s->pool = std::make_unique<Pool>(threads);
  s->force.assign(ngroups, {0.0, 0.0});
  s->init([](size_t, size_t, double& r, double& u, double& v) {
    r = 1.0;
    u = 0.0;
    v = 0.0;
  });
  *out = s.release();
  return VL_OK;
}

// src/cases/mod.rs:11-18
double ramp(uint64_t t, uint64_t steps) {
  if (steps == 0 || t >= steps) return 1.0;
  double s = (double)t / (double)steps;
  return s * s * (3.0 - 2.0 * s);
}

// ---- src/signal.rs -------------------------------------------------------------------------
const double PI = 3.14159265358979323846;

void fft(double* re, double* im, size_t n) {  // signal.rs:6-43
  size_t j = 0;
  for (size_t i = 1; i < n; ++i) {
    size_t bit = n >> 1;
    while (j & bit) {
      j ^= bit;
      bit >>= 1;
    }
    j |= bit;
    if (i < j) {
      std::swap(re[i], re[j]);
      std::swap(im[i], im[j]);
    }
  }
  for (size_t len = 2; len <= n; len <<= 1) {
    double ang = -2.0 * PI / (double)len;
    double wr = std::cos(ang), wi = std::sin(ang);
    for (size_t start = 0; start < n; start += len) {
      double cr = 1.0, ci = 0.0;
      for (size_t k = 0; k < len / 2; ++k) {
        size_t a = start + k, b = start + k + len / 2;
        double tr = re[b] * cr - im[b] * ci, ti = re[b] * ci + im[b] * cr;
        re[b] = re[a] - tr;
        im[b] = im[a] - ti;
        re[a] += tr;
        im[a] += ti;
        double ncr = cr * wr - ci * wi;
        ci = cr * wi + ci * wr;
        cr = ncr;
      }
    }
  }
}

double mean(const double* x, size_t n) {  // signal.rs:56-58
  double s = 0.0;
  for (size_t i = 0; i < n; ++i) s += x[i];
  return s / (double)n;
}

double rms(const double* x, size_t n) {  // signal.rs:61-64
  double m = mean(x, n);
  double s = 0.0;
  for (size_t i = 0; i < n; ++i) s += (x[i] - m) * (x[i] - m);
  return std::sqrt(s / (double)n);
}

size_t next_pow2(size_t v) {
  size_t p = 1;
  while (p < v) p <<= 1;
  return p;
}

// signal.rs:69-88
std::vector<std::array<double, 2>> spectrum(const double* x, size_t n, double dt, size_t pad) {
  double m = mean(x, n);
  size_t size = next_pow2(n * std::max<size_t>(pad, 1));
  std::vector<double> re(size, 0.0), im(size, 0.0);
  for (size_t i = 0; i < n; ++i) {
    double w = 0.5 - 0.5 * std::cos(2.0 * PI * (double)i / (double)n);
    re[i] = (x[i] - m) * w;
  }
  fft(re.data(), im.data(), size);
  double scale = 4.0 / (double)n;
  std::vector<std::array<double, 2>> out(size / 2);
  for (size_t k = 0; k < size / 2; ++k)
    out[k] = {(double)k / ((double)size * dt), scale * std::sqrt(re[k] * re[k] + im[k] * im[k])};
  return out;
}

// signal.rs:92-107
std::array<double, 2> dominant_frequency(const double* x, size_t n, double dt) {
  auto s = spectrum(x, n, dt, 8);
  size_t k = 1;
  for (size_t i = 1; i + 1 < s.size(); ++i)
    if (s[i][1] > s[k][1]) k = i;
  if (k == 0 || k + 1 >= s.size() || s[k][1] <= 0.0 || s[k - 1][1] <= 0.0 || s[k + 1][1] <= 0.0)
    return {s[k][0], s[k][1]};
  double a = std::log(s[k - 1][1]), b = std::log(s[k][1]), c = std::log(s[k + 1][1]);
  double shift = 0.5 * (a - c) / (a - 2.0 * b + c);
  double df = s[1][0] - s[0][0];
  return {s[k][0] + shift * df, std::exp(b - 0.25 * (a - c) * shift)};
}

// signal.rs:109-123
std::vector<std::array<double, 2>> peaks(const double* x, size_t n) {
  std::vector<std::array<double, 2>> out;
  if (n < 3) return out;
  const double m = mean(x, n);
  for (size_t i = 1; i < n - 1; ++i) {
    if (x[i] > m && x[i] > x[i - 1] && x[i] >= x[i + 1]) {
      double a = x[i - 1], b = x[i], c = x[i + 1];
      double denom = a - 2.0 * b + c;
      double shift = denom != 0.0 ? 0.5 * (a - c) / denom : 0.0;
      out.push_back({(double)i + shift, b - 0.25 * (a - c) * shift});
    }
  }
  return out;
}

// signal.rs:127-140
bool crossing_frequency(const double* x, size_t n, double dt, double& f) {
  double m = mean(x, n);
  std::vector<double> times;
  for (size_t i = 1; i < n; ++i) {
    double a = x[i - 1] - m, b = x[i] - m;
    if (a < 0.0 && b >= 0.0) times.push_back((double)(i - 1) + a / (a - b));
  }
  if (times.size() < 2) return false;
  f = (double)(times.size() - 1) / ((times.back() - times.front()) * dt);
  return true;
}

}  // namespace

struct or_channel {  // channel.rs:78-86
  or_sim* sim = nullptr;
  uint64_t ramp_steps = 0;
  bool has_kick = false;
  vl_kick kick{};
  double u_ref = 0.0;
  bool absorbing_outlet = true;
  ~or_channel() { delete sim; }
  void step() {  // channel.rs:136-148
    uint64_t t = sim->t;
    sim->vscale[1] = ramp(t, ramp_steps);
    sim->has_target = absorbing_outlet;
    sim->target = u_ref * sim->vscale[1];
    if (has_kick && t >= kick.start && t < kick.start + kick.duration) {
      double s = (double)(t - kick.start) / (double)kick.duration;
      double sn = std::sin(PI * s);
      double sg = std::sin(2.0 * PI * s);
      double sgn = std::isnan(sg) ? sg : std::copysign(1.0, sg);
      sim->vscale[2] = kick.omega * (sn * sn) * sgn;
    } else {
      sim->vscale[2] = 0.0;
    }
    sim->step();
  }
  void coefficients(double d, double& cd, double& cl) const {  // channel.rs:151-154
    double k = 2.0 / (u_ref * u_ref * d);
    cd = k * sim->force[0][0];
    cl = k * sim->force[0][1];
  }
};

// =====================================================================================
// C API
// =====================================================================================
extern "C" {

const char* or_last_error(void) { return g_err.c_str(); }

static or_shape* wrap_shape(Shape* s) {
  auto* o = new or_shape;
  o->s.reset(s);
  return o;
}

or_shape* or_shape_circle(double cx, double cy, double r) {
  auto* c = new Circle;
  c->cx = cx;
  c->cy = cy;
  c->r = r;
  return wrap_shape(c);
}
or_shape* or_shape_polygon(const double* xy, size_t npts) {
  if (!xy || npts < 3) return nullptr;
  auto* p = new Polygon;
  for (size_t i = 0; i < npts; ++i) p->pts.push_back({xy[2 * i], xy[2 * i + 1]});
  return wrap_shape(p);
}
or_shape* or_shape_rounded_square(double cx, double cy, double half, double r) {
  auto* s = new RoundedSquare;
  s->cx = cx;
  s->cy = cy;
  s->half = half;
  s->r = r;
  return wrap_shape(s);
}
or_shape* or_shape_rotated(or_shape* inner, double cx, double cy, double angle) {
  if (!inner) return nullptr;
  auto* r = new Rotated;
  r->inner = std::move(inner->s);
  delete inner;
  r->cx = cx;
  r->cy = cy;
  r->angle = angle;
  return wrap_shape(r);
}
or_shape* or_shape_translated(or_shape* inner, double dx, double dy) {
  if (!inner) return nullptr;
  auto* t = new Translated;
  t->inner = std::move(inner->s);
  delete inner;
  t->dx = dx;
  t->dy = dy;
  return wrap_shape(t);
}
// geometry.rs:251-261
or_shape* or_shape_section(int corner, double b, double cx, double cy, double d, double angle) {
  or_shape* base;
  if (corner == 2)
    base = or_shape_rounded_square(cx, cy, 0.5 * d, b * d);
  else {
    auto* p = new Polygon;
    p->pts = section_polygon(corner, b, cx, cy, d);
    base = wrap_shape(p);
  }
  if (angle == 0.0) return base;
  return or_shape_rotated(base, cx, cy, angle);
}
or_shape* or_shape_outside_band(double y0, double y1) {
  auto* s = new OutsideBand;
  s->y0 = y0;
  s->y1 = y1;
  return wrap_shape(s);
}
or_shape* or_shape_callback(vl_inside_fn fn, void* ctx) {
  if (!fn) return nullptr;
  auto* c = new Callback;
  c->fn = fn;
  c->ctx = ctx;
  return wrap_shape(c);
}
int or_shape_inside(const or_shape* s, double x, double y) { return s && s->s->inside(x, y) ? 1 : 0; }
double or_cut_fraction(const or_shape* s, double x, double y, double dx, double dy) {
  return cut_fraction(*s->s, x, y, dx, dy);
}
void or_shape_free(or_shape* s) { delete s; }

or_builder* or_builder_new(size_t nx, size_t ny) {
  if (nx < 3 || ny < 3) {
    g_err = "grid must be at least 3 x 3";
    return nullptr;
  }
  auto* b = new or_builder;
  b->nx = nx;
  b->ny = ny;
  b->kinds.emplace_back();  // kinds[0] = Fluid (sim.rs:94)
  b->cell.assign(nx * ny, 0);
  return b;
}
vl_status or_builder_config(or_builder* b, int px, int py, double tau, int collision, double magic,
                            double ax, double ay) {
  if (!b) return fail(VL_EINVAL, "null builder");
  b->periodic_x = px != 0;
  b->periodic_y = py != 0;
  b->tau = tau;
  b->collision = collision;
  b->magic = magic;
  b->accel[0] = ax;
  b->accel[1] = ay;
  return VL_OK;
}
static int add_kind(or_builder* b, Kind&& k) {  // sim.rs:100-104
  if (b->kinds.size() >= 255) {
    g_err = "too many cell kinds";
    return -1;
  }
  b->kinds.push_back(std::move(k));
  return (int)b->kinds.size() - 1;
}
int or_builder_add_wall(or_builder* b, const vl_velocity* vel, int vgroup, int fgroup, or_shape* shape) {
  if (!b) return -1;
  Kind k;
  k.tag = K_WALL;
  k.velocity = make_velocity(vel);
  k.vgroup = vgroup;
  k.fgroup = fgroup;
  if (shape) {
    k.shape = std::move(shape->s);
    delete shape;
  }
  return add_kind(b, std::move(k));
}
int or_builder_add_pressure(or_builder* b, double rho) {
  Kind k;
  k.tag = K_PRESSURE;
  k.rho = rho;
  return add_kind(b, std::move(k));
}
int or_builder_add_radiating(or_builder* b, double rho) {
  Kind k;
  k.tag = K_RADIATING;
  k.rho = rho;
  return add_kind(b, std::move(k));
}
vl_status or_builder_fill_rect(or_builder* b, size_t x0, size_t x1, size_t y0, size_t y1, int kind) {
  if (!b || kind < 0 || (size_t)kind >= b->kinds.size() || x1 >= b->nx || y1 >= b->ny)
    return fail(VL_EINVAL, "fill_rect out of range");
  for (size_t y = y0; y <= y1; ++y)
    for (size_t x = x0; x <= x1; ++x) b->cell[y * b->nx + x] = (uint8_t)kind;
  return VL_OK;
}
vl_status or_builder_fill_shape(or_builder* b, int kind) {  // sim.rs:121-132
  if (!b || kind <= 0 || (size_t)kind >= b->kinds.size() || b->kinds[kind].tag != K_WALL ||
      !b->kinds[kind].shape)
    return fail(VL_EINVAL, "fill_shape needs a wall kind with an outline");
  const Shape& s = *b->kinds[kind].shape;
  for (size_t y = 0; y < b->ny; ++y)
    for (size_t x = 0; x < b->nx; ++x)
      if (s.inside((double)x, (double)y)) b->cell[y * b->nx + x] = (uint8_t)kind;
  return VL_OK;
}
void or_builder_free(or_builder* b) { delete b; }

vl_status or_sim_create(or_builder* b, int threads, or_sim** out) {
  if (!b || !out) return fail(VL_EINVAL, "null argument");
  vl_status st = from_builder(b, threads, out);
  delete b;
  return st;
}
void or_sim_destroy(or_sim* s) { delete s; }
vl_status or_sim_info_get(or_sim* s, vl_sim_info* o) {
  std::memset(o, 0, sizeof(*o));
  o->nx = (int64_t)s->nx;
  o->ny = (int64_t)s->ny;
  o->fluid_nodes = (int64_t)s->fluid_nodes;
  o->link_count = (int64_t)s->links.size();
  o->bnode_count = (int64_t)s->bnodes.size();
  o->pnode_count = (int64_t)s->pnodes.size();
  o->force_groups = (int64_t)s->force.size();
  o->t = (int64_t)s->t;
  o->precision = VL_F64;
  o->collision = s->collision;
  o->omega_plus = s->op;
  o->omega_minus = s->om;
  o->outlet_density_offset = s->outlet_density_offset;
  return VL_OK;
}
vl_status or_sim_init_macro(or_sim* s, const double* rho, const double* ux, const double* uy) {
  s->init([&](size_t x, size_t y, double& r, double& u, double& v) {
    size_t idx = y * s->nx + x;
    r = rho[idx];
    u = ux[idx];
    v = uy[idx];
  });
  return VL_OK;
}
vl_status or_sim_set_vscale(or_sim* s, const double* v) {
  for (int g = 0; g < VGROUPS; ++g) s->vscale[g] = v[g];
  return VL_OK;
}
vl_status or_sim_set_outflow_target(or_sim* s, int enabled, double target) {
  s->has_target = enabled != 0;
  s->target = target;
  return VL_OK;
}
vl_status or_sim_step(or_sim* s, int64_t nsteps) {
  for (int64_t k = 0; k < nsteps; ++k) s->step();
  return VL_OK;
}
vl_status or_sim_run(or_sim* s, int64_t nsteps, const vl_ctl* ctl, double* hist) {
  size_t ng = s->force.size();
  for (int64_t k = 0; k < nsteps; ++k) {
    for (int g = 0; g < VGROUPS; ++g) s->vscale[g] = ctl[k].vscale[g];
    s->has_target = ctl[k].has_target != 0;
    s->target = ctl[k].target;
    s->step();
    if (hist)
      for (size_t g = 0; g < ng; ++g) {
        hist[(k * ng + g) * 2] = s->force[g][0];
        hist[(k * ng + g) * 2 + 1] = s->force[g][1];
      }
  }
  return VL_OK;
}
vl_status or_sim_get_force(or_sim* s, double* out) {
  for (size_t g = 0; g < s->force.size(); ++g) {
    out[2 * g] = s->force[g][0];
    out[2 * g + 1] = s->force[g][1];
  }
  return VL_OK;
}
vl_status or_sim_read(or_sim* s, double* f) {
  std::memcpy(f, s->a.data(), sizeof(double) * s->a.size());
  return VL_OK;
}
vl_status or_sim_write(or_sim* s, const double* f) {
  std::memcpy(s->a.data(), f, sizeof(double) * s->a.size());
  return VL_OK;
}
vl_status or_sim_macros(or_sim* s, double* rho, double* ux, double* uy) {  // sim.rs:527-534
  for (size_t idx = 0; idx < s->n; ++idx) {
    if (s->flag[idx] == 2) {
      rho[idx] = 1.0;
      ux[idx] = 0.0;
      uy[idx] = 0.0;
      continue;
    }
    double r, jx, jy;
    moments(s->pops(idx), r, jx, jy);
    rho[idx] = r;
    ux[idx] = jx / r - 0.5 * s->accel[0];
    uy[idx] = jy / r - 0.5 * s->accel[1];
  }
  return VL_OK;
}
vl_status or_sim_state_hash(or_sim* s, uint64_t* out) {  // sim.rs:654-668
  uint64_t h = 0xcbf29ce484222325ull;
  for (size_t idx = 0; idx < s->n; ++idx) {
    if (s->flag[idx] == 2) continue;
    for (int i = 0; i < Q; ++i) {
      uint64_t bits;
      std::memcpy(&bits, &s->a[i * s->n + idx], 8);
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
struct Neumaier {  // sim.rs:676-691
  double sum = 0.0, comp = 0.0;
  void add(double x) {
    double t = sum + x;
    comp += std::fabs(sum) >= std::fabs(x) ? (sum - t) + x : (x - t) + sum;
    sum = t;
  }
  double value() const { return sum + comp; }
};
}  // namespace
vl_status or_sim_total_mass(or_sim* s, double* out) {  // sim.rs:538-548
  Neumaier m;
  for (size_t idx = 0; idx < s->n; ++idx)
    if (s->flag[idx] != 2)
      for (int i = 0; i < Q; ++i) m.add(s->a[i * s->n + idx]);
  *out = m.value();
  return VL_OK;
}
vl_status or_sim_total_momentum(or_sim* s, double* out) {  // sim.rs:551-563
  Neumaier px, py;
  for (size_t idx = 0; idx < s->n; ++idx)
    if (s->flag[idx] != 2)
      for (int i = 1; i < Q; ++i) {
        double f = s->a[i * s->n + idx];
        px.add((double)CX[i] * f);
        py.add((double)CY[i] * f);
      }
  out[0] = px.value();
  out[1] = py.value();
  return VL_OK;
}
vl_status or_sim_max_speed(or_sim* s, double* out) {  // sim.rs:566-579
  double m = 0.0;
  for (size_t idx = 0; idx < s->n; ++idx)
    if (s->flag[idx] != 2) {
      double ux, uy;
      s->velocity_at(idx, ux, uy);
      double sp = std::sqrt(ux * ux + uy * uy);
      if (std::isnan(sp)) {
        *out = NAN;
        return VL_OK;
      }
      m = std::max(m, sp);
    }
  *out = m;
  return VL_OK;
}

vl_status or_sim_vorticity(or_sim* s, float* out) {  // sim.rs:581-620
  const long nx = (long)s->nx, ny = (long)s->ny;
  std::vector<std::array<double, 2>> u(s->n, {0.0, 0.0});
  for (size_t idx = 0; idx < s->n; ++idx)
    if (s->flag[idx] != 2) s->velocity_at(idx, u[idx][0], u[idx][1]);
  auto fluid = [&](long x, long y) { return x >= 0 && y >= 0 && x < nx && y < ny && s->flag[y * nx + x] != 2; };
  for (long y = 0; y < ny; ++y)
    for (long x = 0; x < nx; ++x) {
      out[y * nx + x] = 0.0f;
      if (!fluid(x, y)) continue;
      auto diff = [&](long dx, long dy, int comp) {
        auto get = [&](long xx, long yy) { return u[yy * nx + xx][comp]; };
        const bool p = fluid(x + dx, y + dy), m = fluid(x - dx, y - dy);
        if (p && m) return 0.5 * (get(x + dx, y + dy) - get(x - dx, y - dy));
        if (p) return get(x + dx, y + dy) - get(x, y);
        if (m) return get(x, y) - get(x - dx, y - dy);
        return 0.0;
      };
      out[y * nx + x] = (float)(diff(1, 0, 1) - diff(0, 1, 0));
    }
  return VL_OK;
}
vl_status or_sim_is_fluid(or_sim* s, int64_t x, int64_t y, int* out) {  // sim.rs:381-383
  if (x < 0 || y < 0 || x >= (int64_t)s->nx || y >= (int64_t)s->ny) return fail(VL_EINVAL, "node outside the domain");
  *out = s->flag[y * s->nx + x] != 2;
  return VL_OK;
}
vl_status or_sim_periodic(or_sim* s, int* px, int* py) {  // sim.rs:670-672
  *px = s->periodic_x;
  *py = s->periodic_y;
  return VL_OK;
}

// channel.rs:89-134
vl_status or_channel_create(const vl_channel_spec* sp, or_shape* body, int threads, or_channel** out) {
  if (!sp || !body || !out) return fail(VL_EINVAL, "null argument");
  size_t nx = (size_t)sp->nx, ny = (size_t)sp->ny;
  double h = (double)(ny - 2);
  or_builder* b = or_builder_new(nx, ny);
  if (!b) {
    or_shape_free(body);
    return VL_EINVAL;
  }
  b->tau = sp->nu / CS2 + 0.5;  // lattice.rs:53-55
  b->collision = sp->collision;
  b->magic = sp->magic;
  vl_velocity inlet_v{};
  if (sp->inflow == VL_INFLOW_PARABOLIC) {
    inlet_v.kind = VL_VEL_PARABOLA;
    inlet_v.p[0] = sp->u;
    inlet_v.p[1] = h;
  } else {
    inlet_v.kind = VL_VEL_CONST;
    inlet_v.p[0] = sp->u;
    inlet_v.p[1] = 0.0;
  }
  int inlet = or_builder_add_wall(b, &inlet_v, 1, -1, nullptr);
  int outlet = or_builder_add_pressure(b, 1.0);
  int walls;
  if (sp->inflow == VL_INFLOW_PARABOLIC) {
    walls = or_builder_add_wall(b, nullptr, 0, -1, nullptr);
  } else if (sp->inflow == VL_INFLOW_UNIFORM) {
    vl_velocity wv{};
    wv.kind = VL_VEL_CONST;
    wv.p[0] = sp->u;
    walls = or_builder_add_wall(b, &wv, 1, -1, nullptr);
  } else {
    walls = or_builder_add_radiating(b, 1.0);
  }
  or_builder_fill_rect(b, 0, 0, 0, ny - 1, inlet);
  or_builder_fill_rect(b, nx - 1, nx - 1, 0, ny - 1, outlet);
  or_builder_fill_rect(b, 1, nx - 2, 0, 0, walls);
  or_builder_fill_rect(b, 1, nx - 2, ny - 1, ny - 1, walls);
  if (sp->inflow == VL_INFLOW_PARABOLIC || sp->inflow == VL_INFLOW_UNIFORM) {
    or_builder_fill_rect(b, 0, nx - 1, 0, 0, walls);
    or_builder_fill_rect(b, 0, nx - 1, ny - 1, ny - 1, walls);
  }
  vl_velocity spin{};
  if (sp->has_kick) {
    spin.kind = VL_VEL_SPIN;
    spin.p[0] = sp->kick.cx;
    spin.p[1] = sp->kick.cy;
  }
  int bodyk = or_builder_add_wall(b, sp->has_kick ? &spin : nullptr, 2, 0, body);
  or_builder_fill_shape(b, bodyk);
  or_sim* sim = nullptr;
  vl_status st = or_sim_create(b, threads, &sim);
  if (st != VL_OK) return st;
  if (sp->ramp_steps == 0) {
    VelFn prof = make_velocity(&inlet_v);
    sim->init([&](size_t, size_t y, double& r, double& u, double& v) {
      r = 1.0;
      u = prof(0.0, (double)y)[0];
      v = 0.0;
    });
  }
  auto* c = new or_channel;
  c->sim = sim;
  c->ramp_steps = sp->ramp_steps;
  c->has_kick = sp->has_kick != 0;
  c->kick = sp->kick;
  c->u_ref = sp->u;
  *out = c;
  return VL_OK;
}
void or_channel_destroy(or_channel* c) { delete c; }
or_sim* or_channel_sim(or_channel* c) { return c->sim; }
vl_status or_channel_set_absorbing(or_channel* c, int a) {
  c->absorbing_outlet = a != 0;
  return VL_OK;
}
vl_status or_channel_run(or_channel* c, int64_t nsteps, double d, double* coef) {
  for (int64_t k = 0; k < nsteps; ++k) {
    c->step();
    double cd = 0.0, cl = 0.0;
    if (!c->sim->force.empty()) c->coefficients(d, cd, cl);
    if (coef) {
      coef[2 * k] = cd;
      coef[2 * k + 1] = cl;
    }
    if (!std::isfinite(cd))
      return fail(VL_ENONFINITE, "diverged after " + std::to_string(c->sim->t) + " steps");
  }
  return VL_OK;
}
vl_status or_channel_coefficients(or_channel* c, double d, double* cd, double* cl) {
  if (c->sim->force.empty()) return fail(VL_EINVAL, "no force group");
  c->coefficients(d, *cd, *cl);
  return VL_OK;
}

// channel.rs:182-206
vl_status or_series_stats(const double* cd_all, const double* cl_all, size_t len, double dt, double from,
                          vl_stats* out) {
  size_t cap = len >= 2 ? len - 2 : 0;
  size_t i0 = std::min((size_t)(from / dt), cap);
  const double* cd = cd_all + i0;
  const double* cl = cl_all + i0;
  size_t n = len - i0;
  double cl_rms = rms(cl, n);
  double st = 0.0, stc = 0.0;
  bool periodic = false;
  if (cl_rms > 1e-6) {
    double f = dominant_frequency(cl, n, dt)[0];
    double fc;
    stc = crossing_frequency(cl, n, dt, fc) ? fc : 0.0;
    periodic = f * (double)n * dt >= 3.0 && std::fabs(stc / f - 1.0) < 0.05;
    if (periodic) st = f;
  }
  out->cd_mean = mean(cd, n);
  out->cd_rms = rms(cd, n);
  out->cl_mean = mean(cl, n);
  out->cl_rms = cl_rms;
  out->st = st;
  out->st_crossings = stc;
  out->periodic = periodic ? 1 : 0;
  return VL_OK;
}
vl_status or_dominant_frequency(const double* x, size_t n, double dt, double* out2) {
  auto r = dominant_frequency(x, n, dt);
  out2[0] = r[0];
  out2[1] = r[1];
  return VL_OK;
}
vl_status or_peaks(const double* x, size_t n, double* pos, double* val, size_t cap, size_t* count) {
  auto p = peaks(x, n);
  for (size_t k = 0; k < p.size() && k < cap; ++k) {
    pos[k] = p[k][0];
    val[k] = p[k][1];
  }
  *count = p.size();
  return VL_OK;
}
vl_status or_crossing_frequency(const double* x, size_t n, double dt, double* out, int* found) {
  double f = 0.0;
  *found = crossing_frequency(x, n, dt, f) ? 1 : 0;
  *out = *found ? f : 0.0;
  return VL_OK;
}
vl_status or_fft(double* re, double* im, size_t n) {
  if (n == 0 || (n & (n - 1))) return fail(VL_EINVAL, "FFT length must be a power of two");
  fft(re, im, n);
  return VL_OK;
}
// signal.rs:47-59 (the O(n^2) transform the FFT is tested against)
vl_status or_dft_naive(const double* re, const double* im, size_t n, double* ore, double* oim) {
  for (size_t k = 0; k < n; ++k) {
    ore[k] = 0.0;
    oim[k] = 0.0;
    for (size_t t = 0; t < n; ++t) {
      double ang = -2.0 * PI * (double)(k * t % n) / (double)n;
      double c = std::cos(ang), s = std::sin(ang);
      ore[k] += re[t] * c - im[t] * s;
      oim[k] += re[t] * s + im[t] * c;
    }
  }
  return VL_OK;
}

// cylinder.rs:35-50 (rig) + :224-232 (bench) with main.rs:448-476's settings.
vl_status or_cylinder_bench(int n, int64_t steps, int threads, uint64_t* hash, int64_t* fluid,
                            int64_t* nx, int64_t* ny, double* seconds) {
  if (n < 10 || n % 10 != 0) return fail(VL_EINVAL, "cells per diameter must be a multiple of 10");
  double nn = (double)n, u_mean = 0.05, re = 100.0, ramp_units = 8.0;
  double xc = 0.5 + 2.0 * nn, yc = 0.5 + 2.0 * nn;
  vl_channel_spec sp{};
  sp.nx = 22 * n + 2;
  sp.ny = 41 * n / 10 + 2;
  sp.inflow = VL_INFLOW_PARABOLIC;
  sp.u = u_mean;
  sp.nu = u_mean * nn / re;
  sp.collision = VL_TRT;
  sp.magic = 3.0 / 16.0;
  sp.ramp_steps = (uint64_t)(ramp_units * nn / u_mean);
  sp.has_kick = 0;
  or_channel* ch = nullptr;
  vl_status st = or_channel_create(&sp, or_shape_circle(xc, yc, 0.5 * nn), threads, &ch);
  if (st != VL_OK) return st;
  auto t0 = std::chrono::steady_clock::now();
  for (int64_t k = 0; k < steps; ++k) ch->step();
  auto t1 = std::chrono::steady_clock::now();
  *seconds = std::chrono::duration<double>(t1 - t0).count();
  or_sim_state_hash(ch->sim, hash);
  *fluid = (int64_t)ch->sim->fluid_nodes;
  *nx = (int64_t)ch->sim->nx;
  *ny = (int64_t)ch->sim->ny;
  delete ch;
  return VL_OK;
}

// poiseuille.rs:41-89
vl_status or_poiseuille(int rows, double tau, int collision, double magic, double q, int threads,
                        double* l2, double* force_balance, int64_t* steps) {
  size_t nx = 4, ny = (size_t)rows + 2;
  double y0 = 1.0 - q, y1 = (double)rows + q;
  double width = y1 - y0;
  double nu = (tau - 0.5) * CS2;  // lattice.rs:48-50
  double umax = 0.1 * 16.0 / (double)rows;
  double g = 8.0 * nu * umax / (width * width);
  or_builder* b = or_builder_new(nx, ny);
  or_builder_config(b, 1, 0, tau, collision, magic, g, 0.0);
  or_shape* shape = (q == 0.5) ? nullptr : or_shape_outside_band(y0, y1);
  int wall = or_builder_add_wall(b, nullptr, 0, 0, shape);
  or_builder_fill_rect(b, 0, nx - 1, 0, 0, wall);
  or_builder_fill_rect(b, 0, nx - 1, ny - 1, ny - 1, wall);
  or_sim* sim = nullptr;
  vl_status st = or_sim_create(b, threads, &sim);
  if (st != VL_OK) return st;
  auto exact = [&](size_t y) { return g / (2.0 * nu) * ((double)y - y0) * (y1 - (double)y); };
  std::vector<double> prev(ny, 0.0);
  auto ux_at = [&](size_t x, size_t y) {
    double ux, uy;
    sim->velocity_at(y * nx + x, ux, uy);
    return ux;
  };
  for (;;) {
    for (int k = 0; k < 200; ++k) sim->step();
    double change = 0.0;
    for (size_t y = 1; y <= (size_t)rows; ++y) {
      double u = ux_at(1, y);
      change = std::max(change, std::fabs(u - prev[y]));
      prev[y] = u;
    }
    if (change < 1e-13 * umax || sim->t > 4000000) break;
  }
  double num = 0.0, den = 0.0;
  for (size_t y = 1; y <= (size_t)rows; ++y) {
    double u = ux_at(1, y), e = exact(y);
    num += (u - e) * (u - e);
    den += e * e;
  }
  double mass;
  or_sim_total_mass(sim, &mass);
  *l2 = std::sqrt(num / den);
  *force_balance = sim->force[0][0] / (g * mass);
  *steps = (int64_t)sim->t;
  delete sim;
  return VL_OK;
}

}  // extern "C"
