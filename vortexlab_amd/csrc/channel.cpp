// Body-in-a-channel driver (src/cases/channel.rs:40-155) above the vl_sim C ABI.
//
// The reference sets three host scalars before every Sim::step (inflow ramp, body kick,
// outflow target) and reads the force back after it.  Here the per-step controls for a whole
// run are tabulated up front (vl_ctl[]) and the force history comes back in one copy, so a
// run is a single stream of kernel launches with no per-step host round trip.
#include <algorithm>
#include <cmath>
#include <memory>
#include <string>
#include <vector>

#include "setup.h"

struct vl_channel {
  vl_sim* sim = nullptr;
  uint64_t ramp_steps = 0;
  bool has_kick = false;
  vl_kick kick{};
  double u_ref = 0.0;
  bool absorbing = true;
  double vscale[4] = {1.0, 1.0, 1.0, 1.0};
  int64_t force_groups = 0;
};

namespace vl {
double channel_kick(const vl_kick& k, bool has_kick, uint64_t t);
}

namespace {

constexpr double kPi = 3.14159265358979323846;

// Controls of step t (src/cases/channel.rs:136-147).
vl_ctl controls(vl_channel* c, uint64_t t) {
  vl_ctl k{};
  c->vscale[1] = vl::ramp(t, c->ramp_steps);
  c->vscale[2] = vl::channel_kick(c->kick, c->has_kick, t);
  for (int g = 0; g < 4; ++g) k.vscale[g] = c->vscale[g];
  k.has_target = c->absorbing ? 1 : 0;
  k.target = c->u_ref * c->vscale[1];
  return k;
}

}  // namespace

extern "C" {

}  // extern "C"

namespace vl {

// The domain of Channel::new (src/cases/channel.rs:89-125): inlet wall cells at x = 0 moving
// with the inflow profile (velocity group 1), pressure outlet at x = nx-1, side walls (at rest,
// moving with the stream, or radiating), and the body (velocity group 2, force group 0).
// `body` is consumed.  Returns the builder, or null with the error set.
vl_builder* channel_builder(const vl_channel_spec* sp, vl_shape* body, vl_velocity* inlet_out) {
  if (!sp || !body) {
    if (body) vl_shape_free(body);
    set_error(VL_EINVAL, "null argument");
    return nullptr;
  }
  if (sp->nx < 3 || sp->ny < 3) {
    vl_shape_free(body);
    set_error(VL_EINVAL, "grid must be at least 3 x 3");
    return nullptr;
  }
  const size_t nx = (size_t)sp->nx, ny = (size_t)sp->ny;
  const double h = (double)(ny - 2);
  vl_builder* b = vl_builder_new(nx, ny);
  vl_builder_config(b, 0, 0, sp->nu / (1.0 / 3.0) + 0.5, sp->collision, sp->magic, 0.0, 0.0);
  vl_velocity inlet_v{};
  if (sp->inflow == VL_INFLOW_PARABOLIC) {
    inlet_v.kind = VL_VEL_PARABOLA;
    inlet_v.p[0] = sp->u;
    inlet_v.p[1] = h;
  } else {
    inlet_v.kind = VL_VEL_CONST;
    inlet_v.p[0] = sp->u;
  }
  if (inlet_out) *inlet_out = inlet_v;
  const int inlet = vl_builder_add_wall(b, &inlet_v, 1, -1, nullptr);
  const int outlet = vl_builder_add_pressure(b, 1.0);
  int walls;
  if (sp->inflow == VL_INFLOW_PARABOLIC) {
    walls = vl_builder_add_wall(b, nullptr, 0, -1, nullptr);
  } else if (sp->inflow == VL_INFLOW_UNIFORM) {
    vl_velocity wv{};
    wv.kind = VL_VEL_CONST;
    wv.p[0] = sp->u;
    walls = vl_builder_add_wall(b, &wv, 1, -1, nullptr);
  } else {
    walls = vl_builder_add_radiating(b, 1.0);
  }
  vl_builder_fill_rect(b, 0, 0, 0, ny - 1, inlet);
  vl_builder_fill_rect(b, nx - 1, nx - 1, 0, ny - 1, outlet);
  vl_builder_fill_rect(b, 1, nx - 2, 0, 0, walls);
  vl_builder_fill_rect(b, 1, nx - 2, ny - 1, ny - 1, walls);
  if (sp->inflow != VL_INFLOW_UNIFORM_OPEN) {  // solid walls also own the corner cells
    vl_builder_fill_rect(b, 0, nx - 1, 0, 0, walls);
    vl_builder_fill_rect(b, 0, nx - 1, ny - 1, ny - 1, walls);
  }
  vl_velocity spin{};
  spin.kind = VL_VEL_SPIN;
  spin.p[0] = sp->kick.cx;
  spin.p[1] = sp->kick.cy;
  const int bodyk = vl_builder_add_wall(b, sp->has_kick ? &spin : nullptr, 2, 0, body);
  if (vl_builder_fill_shape(b, bodyk) != VL_OK) {
    vl_builder_free(b);
    return nullptr;
  }
  return b;
}

// Controls of step t of a channel (src/cases/channel.rs:136-147): the inflow ramp, the body
// kick and the outflow target.
double channel_kick(const vl_kick& k, bool has_kick, uint64_t t) {
  if (!has_kick || t < k.start || t >= k.start + k.duration) return 0.0;
  const double s = (double)(t - k.start) / (double)k.duration;
  const double a = std::sin(kPi * s);
  const double b = std::sin(2.0 * kPi * s);
  const double sgn = std::isnan(b) ? b : std::copysign(1.0, b);  // f64::signum
  return k.omega * (a * a) * sgn;
}

}  // namespace vl

extern "C" {

vl_status vl_channel_create(const vl_channel_spec* sp, vl_shape* body, int precision, vl_channel** out) {
  if (!out) {
    if (body) vl_shape_free(body);
    return vl::set_error(VL_EINVAL, "null argument");
  }
  *out = nullptr;
  vl_velocity inlet_v{};
  vl_builder* b = vl::channel_builder(sp, body, &inlet_v);
  if (!b) return VL_EINVAL;
  const size_t nx = (size_t)sp->nx, ny = (size_t)sp->ny;
  vl_status st;
  vl_sim* sim = nullptr;
  st = vl_sim_create(b, precision, &sim);
  if (st != VL_OK) return st;
  std::unique_ptr<vl_channel> c(new vl_channel);
  c->sim = sim;
  if (sp->ramp_steps == 0) {  // start from the developed inflow profile
    const size_t n = nx * ny;
    std::vector<double> rho(n, 1.0), ux(n, 0.0), uy(n, 0.0);
    const vl::VelocityFn prof = vl::velocity_from_desc(&inlet_v);
    for (size_t y = 0; y < ny; ++y) {
      const double u = prof(0.0, (double)y)[0];
      for (size_t x = 0; x < nx; ++x) ux[y * nx + x] = u;
    }
    st = vl_sim_init_macro(sim, rho.data(), ux.data(), uy.data());
    if (st != VL_OK) {
      vl_sim_destroy(sim);
      return st;
    }
  }
  c->ramp_steps = sp->ramp_steps;
  c->has_kick = sp->has_kick != 0;
  c->kick = sp->kick;
  c->u_ref = sp->u;
  vl_sim_info info;
  vl_sim_info_get(sim, &info);
  c->force_groups = info.force_groups;
  *out = c.release();
  return VL_OK;
}

void vl_channel_destroy(vl_channel* c) {
  if (!c) return;
  vl_sim_destroy(c->sim);
  delete c;
}

vl_sim* vl_channel_sim(vl_channel* c) { return c ? c->sim : nullptr; }

vl_status vl_channel_set_absorbing(vl_channel* c, int absorbing) {
  if (!c) return vl::set_error(VL_EINVAL, "null channel");
  c->absorbing = absorbing != 0;
  return VL_OK;
}

vl_status vl_channel_run(vl_channel* c, int64_t nsteps, double d, double* coef) {
  if (!c) return vl::set_error(VL_EINVAL, "null channel");
  if (nsteps <= 0) return VL_OK;
  vl_sim_info info;
  vl_sim_info_get(c->sim, &info);
  const int64_t ng = c->force_groups;
  const double k = 2.0 / (c->u_ref * c->u_ref * d);  // Channel::coefficients
  // Chunks bound host memory and let a diverged run stop within one chunk of the blow-up.
  const int64_t chunk = 8192;
  std::vector<vl_ctl> ctl;
  std::vector<double> hist;
  for (int64_t k0 = 0; k0 < nsteps; k0 += chunk) {
    const int64_t m = std::min(chunk, nsteps - k0);
    ctl.resize((size_t)m);
    for (int64_t i = 0; i < m; ++i) ctl[(size_t)i] = controls(c, (uint64_t)(info.t + k0 + i));
    hist.assign((size_t)(m * std::max<int64_t>(ng, 1) * 2), 0.0);
    vl_status st = vl_sim_run(c->sim, m, ctl.data(), ng > 0 ? hist.data() : nullptr);
    if (st != VL_OK) return st;
    for (int64_t i = 0; i < m; ++i) {
      const double cd = ng > 0 ? k * hist[(size_t)(i * ng * 2)] : 0.0;
      const double cl = ng > 0 ? k * hist[(size_t)(i * ng * 2 + 1)] : 0.0;
      if (coef) {
        coef[2 * (k0 + i)] = cd;
        coef[2 * (k0 + i) + 1] = cl;
      }
      if (!std::isfinite(cd))
        return vl::set_error(VL_ENONFINITE,
                             "diverged after " + std::to_string(info.t + k0 + i + 1) + " steps (non-finite drag)");
    }
  }
  return VL_OK;
}

vl_status vl_channel_coefficients(vl_channel* c, double d, double* cd, double* cl) {
  if (!c || !cd || !cl) return vl::set_error(VL_EINVAL, "null argument");
  if (c->force_groups < 1) return vl::set_error(VL_EINVAL, "no force group");
  std::vector<double> f((size_t)(2 * c->force_groups));
  vl_status st = vl_sim_get_force(c->sim, f.data());
  if (st != VL_OK) return st;
  const double k = 2.0 / (c->u_ref * c->u_ref * d);
  *cd = k * f[0];
  *cl = k * f[1];
  return VL_OK;
}

}  // extern "C"
