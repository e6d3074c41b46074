// D2Q9 lattice constants and collision operators, usable from host and device code.
//
// Two families:
//  * FP64 "exact" operators on full populations f_i.  They keep the reference's operation
//    order (src/lattice.rs:59-156) so that, compiled with -ffp-contract=off, the GPU path
//    reproduces the reference's populations bit for bit.
//  * FP32 "deviation" operators on d_i = f_i - w_i, the form used by the FP32 and FP16
//    storage paths.  Working on deviations keeps the O(1) rest weight out of every sum, so
//    FP32 resolves the O(Ma) part of the populations to ~1e-7 of itself rather than ~1e-7
//    of w_i.  These contract a*b+c to FMA within each expression only (`contract(on)`): the
//    fusion is then fixed by the source text, so every kernel that inlines them (one step or
//    two steps per pass, single domain or batch) gets the same bits.
#pragma once

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#else
#define __host__
#define __device__
#endif

namespace vl {

constexpr int Q = 9;
// Direction numbering: 0 rest, 1 +x, 2 +y, 3 -x, 4 -y, 5 (+,+), 6 (-,+), 7 (-,-), 8 (+,-)
// (src/lattice.rs:3-20).
__host__ __device__ constexpr int cx(int i) {
  return i == 1 || i == 5 || i == 8 ? 1 : (i == 3 || i == 6 || i == 7 ? -1 : 0);
}
__host__ __device__ constexpr int cy(int i) {
  return i == 2 || i == 5 || i == 6 ? 1 : (i == 4 || i == 7 || i == 8 ? -1 : 0);
}
__host__ __device__ constexpr int opp(int i) { return i == 0 ? 0 : (i <= 4 ? ((i + 1) & 3) + 1 : ((i - 3) & 3) + 5); }
// Weights as correctly rounded FP64 quotients, exactly as the reference's constants.
__host__ __device__ constexpr double wgt(int i) { return i == 0 ? 4.0 / 9.0 : (i <= 4 ? 1.0 / 9.0 : 1.0 / 36.0); }
__host__ __device__ constexpr float wgtf(int i) { return i == 0 ? 4.0f / 9.0f : (i <= 4 ? 1.0f / 9.0f : 1.0f / 36.0f); }
// TRT pairs in the reference's summation order (src/lattice.rs:24).
__host__ __device__ constexpr int pair_a(int p) { return p < 2 ? p + 1 : p + 3; }
__host__ __device__ constexpr int pair_b(int p) { return opp(pair_a(p)); }

static_assert(opp(1) == 3 && opp(2) == 4 && opp(3) == 1 && opp(4) == 2, "opp axis");
static_assert(opp(5) == 7 && opp(6) == 8 && opp(7) == 5 && opp(8) == 6, "opp diagonal");
static_assert(pair_b(0) == 3 && pair_b(1) == 4 && pair_b(2) == 7 && pair_b(3) == 8, "pairs");

enum Model { MODEL_BGK = 0, MODEL_TRT = 1, MODEL_FORCED = 2 };

// ------------------------------ FP64, reference operation order -------------------------
struct Exact {
  __host__ __device__ static inline void moments(const double* f, double& rho, double& jx, double& jy) {
    rho = f[0] + f[1] + f[2] + f[3] + f[4] + f[5] + f[6] + f[7] + f[8];
    jx = f[1] - f[3] + f[5] - f[6] - f[7] + f[8];
    jy = f[2] - f[4] + f[5] + f[6] - f[7] - f[8];
  }

  __host__ __device__ static inline void equilibrium(double rho, double ux, double uy, double* out) {
    double usq = 1.5 * (ux * ux + uy * uy);
#pragma unroll
    for (int i = 0; i < Q; ++i) {
      double cu = 3.0 * ((double)cx(i) * ux + (double)cy(i) * uy);
      out[i] = wgt(i) * rho * (1.0 + cu + 0.5 * cu * cu - usq);
    }
  }

  __host__ __device__ static inline void bgk(double* f, double omega) {
    double rho, jx, jy;
    moments(f, rho, jx, jy);
    double inv = 1.0 / rho;
    double feq[Q];
    equilibrium(rho, jx * inv, jy * inv, feq);
    double moving = 0.0;
#pragma unroll
    for (int i = 1; i < Q; ++i) {
      f[i] = f[i] - omega * (f[i] - feq[i]);
      moving += feq[i];
    }
    f[0] = f[0] - omega * (f[0] - (rho - moving));
  }

  __host__ __device__ static inline void trt(double* f, double op, double om) {
    double rho, jx, jy;
    moments(f, rho, jx, jy);
    double inv = 1.0 / rho;
    double ux = jx * inv, uy = jy * inv;
    double usq = 1.5 * (ux * ux + uy * uy);
    double moving = 0.0;
#pragma unroll
    for (int p = 0; p < 4; ++p) {
      const int a = pair_a(p), b = pair_b(p);
      double cu = 3.0 * ((double)cx(a) * ux + (double)cy(a) * uy);
      double wr = wgt(a) * rho;
      double eq = wr * (1.0 + 0.5 * cu * cu - usq);
      double dp = op * (0.5 * (f[a] + f[b]) - eq);
      double dm = om * (0.5 * (f[a] - f[b]) - wr * cu);
      double fa = f[a], fb = f[b];
      f[a] = fa - dp - dm;
      f[b] = fb - dp + dm;
      moving += 2.0 * eq;
    }
    f[0] = f[0] - op * (f[0] - (rho - moving));
  }

  __host__ __device__ static inline void forced(double* f, double op, double om, double gx, double gy) {
    double rho, jx, jy;
    moments(f, rho, jx, jy);
    double inv = 1.0 / rho;
    double ux = jx * inv + 0.5 * gx, uy = jy * inv + 0.5 * gy;
    double fx = rho * gx, fy = rho * gy;
    double usq = 1.5 * (ux * ux + uy * uy);
    double uf = 3.0 * (ux * fx + uy * fy);
    double kp = 1.0 - 0.5 * op, km = 1.0 - 0.5 * om;
    double moving = 0.0, source = 0.0;
#pragma unroll
    for (int p = 0; p < 4; ++p) {
      const int a = pair_a(p), b = pair_b(p);
      double ccx = (double)cx(a), ccy = (double)cy(a);
      double cu = 3.0 * (ccx * ux + ccy * uy);
      double cf = 3.0 * (ccx * fx + ccy * fy);
      double wr = wgt(a) * rho;
      double eq = wr * (1.0 + 0.5 * cu * cu - usq);
      double dp = op * (0.5 * (f[a] + f[b]) - eq);
      double dm = om * (0.5 * (f[a] - f[b]) - wr * cu);
      double sp = kp * wgt(a) * (cu * cf - uf);
      double sm = km * wgt(a) * cf;
      double fa = f[a], fb = f[b];
      f[a] = fa - dp - dm + sp + sm;
      f[b] = fb - dp + dm + sp - sm;
      moving += 2.0 * eq;
      source += 2.0 * sp;
    }
    f[0] = f[0] - op * (f[0] - (rho - moving)) - source;
  }

  template <int M>
  __host__ __device__ static inline void collide(double* f, double op, double om, double gx, double gy) {
    if (M == MODEL_FORCED)
      forced(f, op, om, gx, gy);
    else if (M == MODEL_BGK)
      bgk(f, op);
    else
      trt(f, op, om);
  }
};

// ------------------------------ FP32 deviation form -------------------------------------
// Device code only (the host translation units built by g++ use the FP64 operators above).
#if defined(__HIPCC__)
// d_i = f_i - w_i.  rho = 1 + sum d_i, j = sum c_i d_i (sum c_i w_i = 0 exactly).
// Equilibrium deviation: feq_i - w_i = w_i (drho + rho (cu + cu^2/2 - 1.5 u^2)).
// The rest population takes the remainder so the deviation mass is conserved in the same
// way as the reference's "rho - moving" rest equilibrium.
struct Dev {
  __device__ static inline void moments(const float* d, float& drho, float& jx, float& jy) {
#pragma clang fp contract(on)
    drho = ((d[0] + d[1]) + (d[2] + d[3])) + ((d[4] + d[5]) + (d[6] + d[7])) + d[8];
    jx = (d[1] - d[3]) + (d[5] - d[6]) + (d[8] - d[7]);
    jy = (d[2] - d[4]) + (d[5] - d[8]) + (d[6] - d[7]);
  }

  // The general form (all three models, one pair at a time); the forced model uses it.
  template <int M>
  __device__ static inline void collide_general(float* d, float op, float om, float gx, float gy) {
#pragma clang fp contract(on)
    float drho, jx, jy;
    moments(d, drho, jx, jy);
    const float rho = 1.0f + drho;
    // the hardware reciprocal (1 ulp): one instruction instead of the ~10 of a correctly rounded
    // FP32 division; every kernel shares this operator, so they still agree bit for bit
    const float inv = __builtin_amdgcn_rcpf(rho);
    float ux = jx * inv, uy = jy * inv;
    float fx = 0.0f, fy = 0.0f, uf = 0.0f, kp = 0.0f, km = 0.0f;
    if (M == MODEL_FORCED) {
      ux += 0.5f * gx;
      uy += 0.5f * gy;
      fx = rho * gx;
      fy = rho * gy;
      uf = 3.0f * (ux * fx + uy * fy);
      kp = 1.0f - 0.5f * op;
      km = 1.0f - 0.5f * om;
    }
    const float usq = 1.5f * (ux * ux + uy * uy);
    float moving = 0.0f, source = 0.0f;
#pragma unroll
    for (int p = 0; p < 4; ++p) {
      const int a = pair_a(p), b = pair_b(p);
      const float w = wgtf(a);
      const float cu = 3.0f * ((float)cx(a) * ux + (float)cy(a) * uy);
      // symmetric equilibrium deviation: w (drho + rho (cu^2/2 - usq))
      const float eqs = w * (drho + rho * (0.5f * cu * cu - usq));
      const float eqa = w * rho * cu;  // antisymmetric equilibrium
      const float fa = d[a], fb = d[b];
      if (M == MODEL_BGK) {
        // BGK: relax each population with op toward eqs +- eqa.
        const float ea = eqs + eqa, eb = eqs - eqa;
        d[a] = fa - op * (fa - ea);
        d[b] = fb - op * (fb - eb);
      } else {
        const float dp = op * (0.5f * (fa + fb) - eqs);
        const float dm = om * (0.5f * (fa - fb) - eqa);
        float na = fa - dp - dm, nb = fb - dp + dm;
        if (M == MODEL_FORCED) {
          const float cf = 3.0f * ((float)cx(a) * fx + (float)cy(a) * fy);
          const float sp = kp * w * (cu * cf - uf);
          const float sm = km * w * cf;
          na += sp + sm;
          nb += sp - sm;
          source += 2.0f * sp;
        }
        d[a] = na;
        d[b] = nb;
      }
      moving += 2.0f * eqs;
    }
    const float eq0 = drho - moving;
    d[0] = d[0] - op * (d[0] - eq0) - source;
  }

  // BGK and TRT in fewer operations (the pass kernels are bound by vector instruction issue):
  // the pair sums s = d_a + d_b and differences a = d_a - d_b give the moments and the TRT
  // symmetric / antisymmetric parts at once; c.u of the four first members is ux, uy, ux + uy,
  // uy - ux (no products with zero components); the equilibrium deviation of a pair,
  //   eqs = w (drho + rho (4.5 (c.u)^2 - 1.5 u^2)) = w drho + 1.5 w rho t,  t = 3 (c.u)^2 - u^2,
  //   eqa = w rho 3 c.u,
  // enters with the relaxation rates folded into its per-weight factors.  The updates keep the
  // difference form (the non-equilibrium part is formed before it is scaled), and the rest
  // population takes the remainder drho - sum of the moving equilibria, so mass is conserved as
  // in the reference's "rho - moving".
  template <int M>
  __device__ static inline void collide(float* d, float op, float om, float gx, float gy) {
    if constexpr (M == MODEL_FORCED) {
      collide_general<M>(d, op, om, gx, gy);
    } else {
#pragma clang fp contract(on)
      constexpr float WA = 1.0f / 9.0f, WD = 1.0f / 36.0f;
      const float s13 = d[1] + d[3], s24 = d[2] + d[4], s57 = d[5] + d[7], s68 = d[6] + d[8];
      const float a13 = d[1] - d[3], a24 = d[2] - d[4], a57 = d[5] - d[7], a68 = d[6] - d[8];
      const float drho = ((d[0] + s13) + (s24 + s57)) + s68;
      const float jx = a13 + (a57 - a68), jy = a24 + (a57 + a68);
      const float rho = 1.0f + drho;
      // the hardware reciprocal (1 ulp): one instruction instead of the ~10 of a correctly
      // rounded FP32 division; every kernel shares this operator, so they agree bit for bit
      const float inv = __builtin_amdgcn_rcpf(rho);
      const float ux = jx * inv, uy = jy * inv;
      const float uu = ux * ux + uy * uy;
      const float sp[4] = {s13, s24, s57, s68}, ap[4] = {a13, a24, a57, a68};
      // c.u of the first member of pair p (1, 2, 5, 6), formed where the pair is updated so that
      // only one pair's terms are live at a time
      auto cu_of = [&](int p) { return p == 0 ? ux : (p == 1 ? uy : (p == 2 ? ux + uy : uy - ux)); };
      if constexpr (M == MODEL_BGK) {
        float moving = 0.0f;
#pragma unroll
        for (int p = 0; p < 4; ++p) {
          const float w = p < 2 ? WA : WD;
          const int ia = pair_a(p), ib = pair_b(p);
          const float c = cu_of(p), k = 3.0f * c, t = k * c - uu;
          const float eqs = (1.5f * w * rho) * t + w * drho;
          const float eqa = (w * rho) * k;
          const float ea = eqs + eqa, eb = eqs - eqa;
          d[ia] = d[ia] - op * (d[ia] - ea);
          d[ib] = d[ib] - op * (d[ib] - eb);
          moving = 2.0f * eqs + moving;
        }
        d[0] = d[0] - op * (d[0] - (drho - moving));
      } else {
        const float hp = 0.5f * op, hm = 0.5f * om;
        float tsum[2];
#pragma unroll
        for (int p = 0; p < 4; ++p) {
          const float w = p < 2 ? WA : WD;
          const int ia = pair_a(p), ib = pair_b(p);
          const float c = cu_of(p), k = 3.0f * c, t = k * c - uu;
          if (p % 2 == 0)
            tsum[p / 2] = t;
          else
            tsum[p / 2] = tsum[p / 2] + t;
          const float ope = ((1.5f * op * w) * rho) * t + (op * w) * drho;  // op * eqs
          const float ome = ((om * w) * rho) * k;                            // om * eqa
          const float dp = hp * sp[p] - ope;                                 // op (s/2 - eqs)
          const float dm = hm * ap[p] - ome;                                 // om (a/2 - eqa)
          d[ia] = d[ia] - dp - dm;
          d[ib] = d[ib] - dp + dm;
        }
        // sum of the moving equilibria: 2 sum eqs = rho (3 WA (t1 + t2) + 3 WD (t5 + t6)) + 4 (WA + WD) drho
        const float m = (3.0f * WA) * tsum[0] + (3.0f * WD) * tsum[1];
        const float moving = rho * m + (4.0f * (WA + WD)) * drho;
        d[0] = d[0] - op * (d[0] - (drho - moving));
      }
    }
  }
};

#endif  // __HIPCC__

}  // namespace vl
