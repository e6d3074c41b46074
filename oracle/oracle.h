/*
 * CPU oracle: a C++ restatement of the reference crate's D2Q9 solver (src/lattice.rs,
 * src/sim.rs, src/geometry.rs, src/cases/{mod,channel,cylinder,tunnel,poiseuille}.rs,
 * src/signal.rs), operation for operation, compiled with -ffp-contract=off.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product (vortexlab_amd/) links, loads or
 * calls this code; only tests/, __graft_entry__.smoke() and the cpu_baseline leg of
 * bench.py use it, as the checker and the timed CPU baseline.
 *
 * The entry points mirror include/vortexlab.h with an `or_` prefix and the same data
 * structs, so one test can drive the GPU library and the oracle with identical calls.
 */
#ifndef VL_ORACLE_H
#define VL_ORACLE_H

#include "../include/vortexlab.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct or_shape or_shape;
typedef struct or_builder or_builder;
typedef struct or_sim or_sim;
typedef struct or_channel or_channel;

const char* or_last_error(void);

or_shape* or_shape_circle(double cx, double cy, double r);
or_shape* or_shape_polygon(const double* xy, size_t npts);
or_shape* or_shape_rounded_square(double cx, double cy, double half, double r);
or_shape* or_shape_rotated(or_shape* inner, double cx, double cy, double angle);
or_shape* or_shape_translated(or_shape* inner, double dx, double dy);
or_shape* or_shape_section(int corner, double b, double cx, double cy, double d, double angle);
or_shape* or_shape_outside_band(double y0, double y1);
or_shape* or_shape_callback(vl_inside_fn fn, void* ctx);
int or_shape_inside(const or_shape* s, double x, double y);
double or_cut_fraction(const or_shape* s, double x, double y, double dx, double dy);
void or_shape_free(or_shape* s);

or_builder* or_builder_new(size_t nx, size_t ny);
vl_status or_builder_config(or_builder* b, int periodic_x, int periodic_y, double tau,
                            int collision, double magic, double ax, double ay);
int or_builder_add_wall(or_builder* b, const vl_velocity* vel, int vgroup, int fgroup,
                        or_shape* shape);
int or_builder_add_pressure(or_builder* b, double rho);
int or_builder_add_radiating(or_builder* b, double rho);
vl_status or_builder_fill_rect(or_builder* b, size_t x0, size_t x1, size_t y0, size_t y1, int kind);
vl_status or_builder_fill_shape(or_builder* b, int kind);
void or_builder_free(or_builder* b);

/* `threads` is the reference's Pool size (static row split, src/sim.rs:401-403). */
vl_status or_sim_create(or_builder* b, int threads, or_sim** out);
void or_sim_destroy(or_sim* s);
vl_status or_sim_info_get(or_sim* s, vl_sim_info* out);
vl_status or_sim_init_macro(or_sim* s, const double* rho, const double* ux, const double* uy);
vl_status or_sim_set_vscale(or_sim* s, const double* vscale4);
vl_status or_sim_set_outflow_target(or_sim* s, int enabled, double target);
vl_status or_sim_step(or_sim* s, int64_t nsteps);
vl_status or_sim_run(or_sim* s, int64_t nsteps, const vl_ctl* ctl, double* force_hist);
vl_status or_sim_get_force(or_sim* s, double* out);
vl_status or_sim_read(or_sim* s, double* f);
vl_status or_sim_write(or_sim* s, const double* f);
vl_status or_sim_macros(or_sim* s, double* rho, double* ux, double* uy);
vl_status or_sim_state_hash(or_sim* s, uint64_t* out);
vl_status or_sim_total_mass(or_sim* s, double* out);
vl_status or_sim_total_momentum(or_sim* s, double* out2);
vl_status or_sim_max_speed(or_sim* s, double* out);
vl_status or_sim_vorticity(or_sim* s, float* out);
vl_status or_sim_is_fluid(or_sim* s, int64_t x, int64_t y, int* out);
vl_status or_sim_periodic(or_sim* s, int* px, int* py);

vl_status or_channel_create(const vl_channel_spec* spec, or_shape* body, int threads,
                            or_channel** out);
void or_channel_destroy(or_channel* c);
or_sim* or_channel_sim(or_channel* c);
vl_status or_channel_set_absorbing(or_channel* c, int absorbing);
vl_status or_channel_run(or_channel* c, int64_t nsteps, double d, double* coef);
vl_status or_channel_coefficients(or_channel* c, double d, double* cd, double* cl);

vl_status or_series_stats(const double* cd, const double* cl, size_t n, double dt,
                          double from, vl_stats* out);
vl_status or_dominant_frequency(const double* x, size_t n, double dt, double* out2);
vl_status or_peaks(const double* x, size_t n, double* pos, double* val, size_t cap, size_t* count);
vl_status or_crossing_frequency(const double* x, size_t n, double dt, double* out, int* found);
vl_status or_fft(double* re, double* im, size_t n);
vl_status or_dft_naive(const double* re, const double* im, size_t n, double* ore, double* oim);

/* cylinder::bench (src/cases/cylinder.rs:224-232) with the CLI's settings
 * (src/main.rs:448-476: u_mean 0.05, TRT 3/16): returns the state hash, fluid-node count,
 * grid and wall-clock seconds of the step loop. */
vl_status or_cylinder_bench(int n, int64_t steps, int threads, uint64_t* hash,
                            int64_t* fluid_nodes, int64_t* nx, int64_t* ny, double* seconds);

/* poiseuille::run (src/cases/poiseuille.rs:41-89): l2 error, force balance, steps. */
vl_status or_poiseuille(int rows, double tau, int collision, double magic, double q,
                        int threads, double* l2, double* force_balance, int64_t* steps);

#ifdef __cplusplus
}
#endif
#endif
