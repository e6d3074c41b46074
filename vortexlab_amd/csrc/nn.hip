// Surrogate network: shape mask (64x64) + log Re -> (mean Cd, Cl', St), trained and evaluated
// by the kernels in this file (no reference counterpart).
//
//   4 blocks: conv 3x3 (pad 1) + bias + ReLU, then 2x2 max-pool
//             channels 1 -> 32 -> 32 -> 64 -> 64, spatial 64 -> 32 -> 16 -> 8 -> 4
//   flatten 64*4*4 = 1024, append the scaled log Re -> 1025
//   FC 1025 -> 128 + ReLU, FC 128 -> 3
//
// FP32 throughout.  Every reduction (over a batch, over pixels) runs in a fixed order with a
// fixed tree, never with atomics, so training is bitwise reproducible from the seed.  The
// dataset is kept on the device and batches are gathered by index on the device.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <random>
#include <string>
#include <vector>

#include "setup.h"
#include "sim_internal.h"

namespace vl {
namespace nn {

constexpr int IMG = 64;
constexpr int NB = 4;                            // conv blocks
constexpr int CH[NB + 1] = {1, 32, 32, 64, 64};  // channels
constexpr int SP[NB + 1] = {64, 32, 16, 8, 4};   // input side of block l (SP[l]), pooled side SP[l+1]
constexpr int FLAT = 64 * 4 * 4;
constexpr int FIN = FLAT + 1;
constexpr int HID = 128;
constexpr int OUT = 3;
constexpr int T = 256;

// parameter offsets
struct Layout {
  int64_t cw[NB], cb[NB], w1, b1, w2, b2, total;
  Layout() {
    int64_t o = 0;
    for (int l = 0; l < NB; ++l) {
      cw[l] = o;
      o += (int64_t)CH[l + 1] * CH[l] * 9;
      cb[l] = o;
      o += CH[l + 1];
    }
    w1 = o;
    o += (int64_t)HID * FIN;
    b1 = o;
    o += HID;
    w2 = o;
    o += (int64_t)OUT * HID;
    b2 = o;
    o += OUT;
    total = o;
  }
};

__global__ void k_gather(const float* __restrict__ masks, const float* __restrict__ feat,
                         const float* __restrict__ targ, const int32_t* __restrict__ idx, int n, float* __restrict__ x0,
                         float* __restrict__ s, float* __restrict__ t) {
  const int64_t total = (int64_t)n * IMG * IMG;
  for (int64_t e = (int64_t)blockIdx.x * T + threadIdx.x; e < total; e += (int64_t)gridDim.x * T) {
    const int64_t i = e / (IMG * IMG), p = e - i * IMG * IMG;
    x0[e] = masks[(int64_t)idx[i] * IMG * IMG + p];
  }
  const int64_t g = (int64_t)blockIdx.x * T + threadIdx.x;
  if (g < n) {
    s[g] = feat[idx[g]];
    if (targ)
      for (int k = 0; k < OUT; ++k) t[g * OUT + k] = targ[(int64_t)idx[g] * OUT + k];
  }
}

// ---- convolutions as implicit GEMMs on the matrix cores (FP32 in, FP32 accumulate) ----------
// One wave computes a 16 x 16 tile D[i][j] = sum_k A[i][k] B[k][j] with v_mfma_f32_16x16x4f32,
// four k at a time: lane l supplies A[i0 + l % 16][k0 + l / 16] and B[k0 + l / 16][j0 + l % 16]
// and receives D[i0 + 4 (l / 16) + r][j0 + l % 16] in accumulator register r.  A tile's k run in
// one fixed order, so every result is reproducible bit for bit.  Pixel tiles are 16 consecutive
// pixels of one sample; out-of-range taps read as 0.
typedef float f4 __attribute__((ext_vector_type(4)));
constexpr int MW = 4;  // waves per workgroup of the matrix-core kernels


// Tap k = ci * 9 + ky * 3 + kx of a 3x3 kernel.
__device__ inline void tap(int k, int& ci, int& ky, int& kx) {
  ci = k / 9;
  const int r = k - ci * 9;
  ky = r / 3;
  kx = r - ky * 3;
}

// 2x2 max-pool: out side = side/2
__global__ void k_pool_fwd(const float* __restrict__ in, float* __restrict__ out, int nc, int side) {
  const int h = side / 2;
  const int64_t total = (int64_t)nc * h * h;
  for (int64_t e = (int64_t)blockIdx.x * T + threadIdx.x; e < total; e += (int64_t)gridDim.x * T) {
    const int x = (int)(e % h), y = (int)((e / h) % h);
    const int64_t c = e / ((int64_t)h * h);
    const float* p = in + (c * side + 2 * y) * side + 2 * x;
    out[e] = fmaxf(fmaxf(p[0], p[1]), fmaxf(p[side], p[side + 1]));
  }
}

// gradient through the pool and the ReLU before it: the first maximum of each window (raster
// order) takes the gradient; zero where the activation is zero.
__global__ void k_pool_bwd(const float* __restrict__ act, const float* __restrict__ dpool, float* __restrict__ dact,
                           int nc, int side) {
  const int64_t total = (int64_t)nc * side * side;
  const int h = side / 2;
  for (int64_t e = (int64_t)blockIdx.x * T + threadIdx.x; e < total; e += (int64_t)gridDim.x * T) {
    const int x = (int)(e % side), y = (int)((e / side) % side);
    const int64_t c = e / ((int64_t)side * side);
    const int px = x / 2, py = y / 2;
    const float* p = act + (c * side + 2 * py) * side + 2 * px;
    const float v[4] = {p[0], p[1], p[side], p[side + 1]};
    int am = 0;
    for (int k = 1; k < 4; ++k)
      if (v[k] > v[am]) am = k;
    const int me = (y - 2 * py) * 2 + (x - 2 * px);
    const float a = act[e];
    dact[e] = (me == am && a > 0.0f) ? dpool[(c * h + py) * h + px] : 0.0f;
  }
}

// The forward conv and the conv's input gradient as one implicit GEMM, D = A B:
// * forward (FWD): out[n][co][p] = relu(b[co] + sum_(ci,ky,kx) w[co][ci][ky][kx] in[n][ci][p + (ky-1, kx-1)]),
//   A = w (co x ci*9), B = the input's taps;
// * input gradient: din[n][ci][p] = sum_(co,ky,kx) w[co][ci][ky][kx] dz[n][co][p - (ky-1, kx-1)],
//   A[ci][(co,ky,kx)] = w, B = dz's taps.
// A wave covers 16 rows of D (output channels) x PT pixel tiles: one weight load and tap decode
// per k serve PT products, and the k loop issues the loads of U steps before their products.
// With `splits` > 1 the k range is cut into fixed pieces of `kper` and each piece's sums go to
// part[piece][...]; k_conv_split_sum then adds the pieces in order.  The pieces depend only on
// the layer, never on the batch, so a sample's result does not depend on its batch.
constexpr int U = 4;
template <int PT, bool FWD>
__global__ __launch_bounds__(MW * 64) void k_conv_mm(const float* __restrict__ src, const float* __restrict__ w,
                                                     const float* __restrict__ b, float* __restrict__ dst,
                                                     float* __restrict__ part, int n, int ci_n, int co_n, int side,
                                                     int splits, int kper) {
  const int lane = (int)(threadIdx.x & 63);
  const int S2 = side * side;
  const int M = FWD ? co_n : ci_n, KC = FWD ? ci_n : co_n, K = KC * 9;
  const int mtiles = M / 16, ptiles = n * S2 / (16 * PT);
  const int t = (int)blockIdx.x * MW + (int)(threadIdx.x >> 6);
  if (t >= ptiles * mtiles * splits) return;
  const int m0 = (t % mtiles) * 16, pt = (t / mtiles) % ptiles, piece = t / (mtiles * ptiles);
  const int smp = pt * 16 * PT / S2, p0 = pt * 16 * PT - smp * S2;
  const int i = lane & 15, kq = lane >> 4;
  int y[PT], x[PT];
#pragma unroll
  for (int q = 0; q < PT; ++q) {
    y[q] = (p0 + 16 * q + i) / side;
    x[q] = (p0 + 16 * q + i) - y[q] * side;
  }
  const float* sbase = src + (int64_t)smp * KC * S2;
  const int kbeg = piece * kper, kend = min(K, kbeg + kper);
  f4 acc[PT];
#pragma unroll
  for (int q = 0; q < PT; ++q) acc[q] = f4{0.0f, 0.0f, 0.0f, 0.0f};
  for (int k0 = kbeg; k0 < kend; k0 += 4 * U) {
    float a[U], v[U][PT];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int k = k0 + 4 * u + kq;
      a[u] = 0.0f;
#pragma unroll
      for (int q = 0; q < PT; ++q) v[u][q] = 0.0f;
      if (k < kend) {
        int c, ky, kx;
        tap(k, c, ky, kx);
        a[u] = FWD ? w[(int64_t)(m0 + i) * K + k] : w[((int64_t)c * ci_n + m0 + i) * 9 + ky * 3 + kx];
        const int dy = FWD ? ky - 1 : 1 - ky, dx = FWD ? kx - 1 : 1 - kx;
        const float* cp = sbase + (int64_t)c * S2;
#pragma unroll
        for (int q = 0; q < PT; ++q) {
          const int yy = y[q] + dy, xx = x[q] + dx;
          if (yy >= 0 && yy < side && xx >= 0 && xx < side) v[u][q] = cp[yy * side + xx];
        }
      }
    }
#pragma unroll
    for (int u = 0; u < U; ++u)
#pragma unroll
      for (int q = 0; q < PT; ++q) acc[q] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[u], v[u][q], acc[q], 0, 0, 0);
  }
  const int64_t total = (int64_t)n * M * S2;
  float* out = splits > 1 ? part + piece * total : dst;
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const int m = m0 + 4 * kq + r;
    float* op = out + ((int64_t)smp * M + m) * S2 + p0 + i;
    const float bc = (FWD && splits == 1) ? b[m] : 0.0f;
#pragma unroll
    for (int q = 0; q < PT; ++q) {
      const float z = acc[q][r] + bc;
      op[16 * q] = (FWD && splits == 1) ? (z > 0.0f ? z : 0.0f) : z;
    }
  }
}

// dst[e] = sum of the pieces' sums in piece order (+ bias and ReLU for the forward conv).
__global__ void k_conv_split_sum(const float* __restrict__ part, const float* __restrict__ b, float* __restrict__ dst,
                                 int64_t total, int M, int S2, int splits, int fwd) {
  for (int64_t e = (int64_t)blockIdx.x * T + threadIdx.x; e < total; e += (int64_t)gridDim.x * T) {
    float z = 0.0f;
    for (int sp = 0; sp < splits; ++sp) z += part[sp * total + e];
    if (fwd) {
      z += b[(int)((e / S2) % M)];
      z = z > 0.0f ? z : 0.0f;
    }
    dst[e] = z;
  }
}

// Weight and bias gradients of a conv, split over the batch-pixel sum in fixed chunks:
// part[c][co][j] = sum over k = (sample, pixel) in chunk c of dz[k][co] * B[k][j], with
// B[k][(ci,ky,kx)] the input's taps and B[k][ci_n * 9] = 1 (the bias).  A[co][k] = dz.  A wave
// covers 16 output channels x JT column tiles (one dz load and pixel decode per k serve JT
// products).  nj = ci_n * 9 + 1 columns, padded to a multiple of 16 JT (`njp`); co_n a multiple
// of 16.
template <int JT>
__global__ __launch_bounds__(MW * 64) void k_conv_bwd_w(const float* __restrict__ dz, const float* __restrict__ in,
                                                        float* __restrict__ part, int n, int ci_n, int co_n, int side,
                                                        int chunk, int nchunks, int njp) {
  const int lane = (int)(threadIdx.x & 63);
  const int S2 = side * side, ctiles = co_n / 16, jgroups = njp / (16 * JT), nj = ci_n * 9 + 1;
  const int t = (int)blockIdx.x * MW + (int)(threadIdx.x >> 6);
  if (t >= nchunks * ctiles * jgroups) return;
  const int jg = t % jgroups, ct = (t / jgroups) % ctiles, c = t / (jgroups * ctiles);
  const int co0 = ct * 16, j0 = jg * 16 * JT;
  const int i = lane & 15, kq = lane >> 4;
  // column g's tap offset into the sample's input planes (-1: the bias column, -2: padding)
  int64_t off[JT];
  int dy[JT], dx[JT];
#pragma unroll
  for (int g = 0; g < JT; ++g) {
    const int j = j0 + 16 * g + i;
    int ci = 0, ky = 1, kx = 1;
    if (j < nj - 1) tap(j, ci, ky, kx);
    off[g] = j < nj - 1 ? (int64_t)ci * S2 : (j == nj - 1 ? -1 : -2);
    dy[g] = ky - 1;
    dx[g] = kx - 1;
  }
  const int64_t kbeg = (int64_t)c * chunk, kend = std::min<int64_t>(kbeg + chunk, (int64_t)n * S2);
  f4 acc[JT];
#pragma unroll
  for (int g = 0; g < JT; ++g) acc[g] = f4{0.0f, 0.0f, 0.0f, 0.0f};
  for (int64_t k0 = kbeg; k0 < kend; k0 += 4 * U) {
    float a[U], v[U][JT];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int64_t k = k0 + 4 * u + kq;  // n * S2 and the chunk are multiples of 4
      a[u] = 0.0f;
#pragma unroll
      for (int g = 0; g < JT; ++g) v[u][g] = 0.0f;
      if (k < kend) {
        const int smp = (int)(k / S2), p = (int)(k - (int64_t)smp * S2);
        const int y = p / side, x = p - y * side;
        a[u] = dz[((int64_t)smp * co_n + co0 + i) * S2 + p];
        const float* ip = in + (int64_t)smp * ci_n * S2;
#pragma unroll
        for (int g = 0; g < JT; ++g) {
          const int yy = y + dy[g], xx = x + dx[g];
          if (off[g] >= 0) {
            if (yy >= 0 && yy < side && xx >= 0 && xx < side) v[u][g] = ip[off[g] + yy * side + xx];
          } else if (off[g] == -1) {
            v[u][g] = 1.0f;
          }
        }
      }
    }
#pragma unroll
    for (int u = 0; u < U; ++u)
#pragma unroll
      for (int g = 0; g < JT; ++g) acc[g] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[u], v[u][g], acc[g], 0, 0, 0);
  }
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    float* pp = part + ((int64_t)c * co_n + co0 + 4 * kq + r) * njp + j0 + i;
#pragma unroll
    for (int g = 0; g < JT; ++g) pp[16 * g] = acc[g][r];
  }
}

// Sums the chunks' partial gradients into dw[co][ci][ky][kx] and db[co]: one wave per output,
// lane l adds chunks l, l + 64, ... in order, then a fixed butterfly over the lanes.
__global__ __launch_bounds__(T) void k_conv_bwd_w_sum(const float* __restrict__ part, float* __restrict__ dw,
                                                      float* __restrict__ db, int ci_n, int co_n, int nchunks,
                                                      int njp) {
  const int nj = ci_n * 9 + 1;
  const int lane = (int)(threadIdx.x & 63);
  const int e = (int)(blockIdx.x * (T / 64) + (threadIdx.x >> 6));
  if (e >= co_n * nj) return;
  const int co = e / nj, j = e - co * nj;
  float acc = 0.0f;
  for (int c = lane; c < nchunks; c += 64) acc += part[((int64_t)c * co_n + co) * njp + j];
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) acc += __shfl_xor(acc, d, 64);
  if (lane == 0) {
    if (j < nj - 1)
      dw[(int64_t)co * (nj - 1) + j] = acc;
    else
      db[co] = acc;
  }
}

// fully connected: out[s][m] = b[m] + sum_k w[m][k] x[s][k] (optional ReLU).  The FC1 input is the
// flattened last pooled map followed by the scalar feature.
__global__ __launch_bounds__(T) void k_fc_fwd(const float* __restrict__ x, const float* __restrict__ xs, int kx,
                                              const float* __restrict__ w, const float* __restrict__ b,
                                              float* __restrict__ out, int n, int m_n, int relu) {
  // one wave per output: lane l sums k = l, l + 64, ... in order, then a fixed butterfly over the lanes
  const int lane = (int)(threadIdx.x & 63);
  const int64_t e = (int64_t)blockIdx.x * (T / 64) + (threadIdx.x >> 6);
  if (e >= (int64_t)n * m_n) return;
  const int kin = kx + (xs ? 1 : 0);
  const int m = (int)(e % m_n);
  const int64_t smp = e / m_n;
  const float* wp = w + (int64_t)m * kin;
  const float* xp = x + smp * kx;
  float acc = 0.0f;
  for (int k = lane; k < kin; k += 64) acc += wp[k] * (k < kx ? xp[k] : xs[smp]);
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) acc += __shfl_xor(acc, d, 64);
  if (lane == 0) {
    const float z = b[m] + acc;
    out[e] = relu ? fmaxf(z, 0.0f) : z;
  }
}

// FC backward: weight / bias gradients summed over the batch in sample order.
__global__ void k_fc_bwd_w(const float* __restrict__ dy, const float* __restrict__ x, const float* __restrict__ xs,
                           int kx, float* __restrict__ dw, float* __restrict__ db, int n, int m_n) {
  const int kin = kx + (xs ? 1 : 0);
  const int64_t total = (int64_t)m_n * (kin + 1);
  for (int64_t e = (int64_t)blockIdx.x * T + threadIdx.x; e < total; e += (int64_t)gridDim.x * T) {
    const int m = (int)(e / (kin + 1));
    const int k = (int)(e - (int64_t)m * (kin + 1));
    float acc = 0.0f;
    for (int s = 0; s < n; ++s) {
      const float v = k < kx ? x[(int64_t)s * kx + k] : (k == kx && xs ? xs[s] : 1.0f);
      acc += dy[(int64_t)s * m_n + m] * v;
    }
    if (k < kin)
      dw[(int64_t)m * kin + k] = acc;
    else
      db[m] = acc;
  }
}

// FC backward: input gradient dx[s][k] = sum_m w[m][k] dy[s][m], times the ReLU mask of x.
__global__ void k_fc_bwd_x(const float* __restrict__ dy, const float* __restrict__ w, const float* __restrict__ xact,
                           int kin, int kx, float* __restrict__ dx, int n, int m_n, int mask_relu) {
  const int64_t total = (int64_t)n * kx;
  for (int64_t e = (int64_t)blockIdx.x * T + threadIdx.x; e < total; e += (int64_t)gridDim.x * T) {
    const int k = (int)(e % kx);
    const int64_t s = e / kx;
    float acc = 0.0f;
    for (int m = 0; m < m_n; ++m) acc += w[(int64_t)m * kin + k] * dy[s * m_n + m];
    if (mask_relu && !(xact[e] > 0.0f)) acc = 0.0f;
    dx[e] = acc;
  }
}

// MSE loss (mean over n x OUT) and its gradient; the loss is a fixed-order tree in one workgroup.
__global__ __launch_bounds__(T) void k_mse(const float* __restrict__ y, const float* __restrict__ t, float* __restrict__ dy,
                                           float* __restrict__ loss, int n) {
  __shared__ float red[T];
  const int total = n * OUT;
  const float inv = 1.0f / (float)total;
  float acc = 0.0f;
  for (int e = threadIdx.x; e < total; e += T) {
    const float d = y[e] - t[e];
    dy[e] = 2.0f * d * inv;
    acc += d * d;
  }
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int w = T / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) *loss = red[0] * inv;
}

__global__ void k_adam(float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m, float* __restrict__ v,
                       int64_t n, float lr, float b1, float b2, float eps, float c1, float c2) {
  for (int64_t e = (int64_t)blockIdx.x * T + threadIdx.x; e < n; e += (int64_t)gridDim.x * T) {
    const float gi = g[e];
    const float mi = b1 * m[e] + (1.0f - b1) * gi;
    const float vi = b2 * v[e] + (1.0f - b2) * gi * gi;
    m[e] = mi;
    v[e] = vi;
    p[e] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
  }
}

}  // namespace nn
}  // namespace vl

using namespace vl;
using namespace vl::nn;

struct vl_nn {
  Layout L;
  int max_batch = 0;
  hipStream_t stream = nullptr;
  float *p = nullptr, *g = nullptr, *m = nullptr, *v = nullptr;
  // activations
  float* x0 = nullptr;           // gathered masks [B][64][64]
  float* s = nullptr;            // scaled log Re [B]
  float* tgt = nullptr;          // targets [B][3]
  float* act[NB] = {};           // post-ReLU conv output [B][C][side][side]
  float* pool[NB] = {};          // pooled [B][C][side/2][side/2]
  float *h = nullptr, *y = nullptr, *dy = nullptr, *dh = nullptr, *dflat = nullptr;
  float* dact = nullptr;         // scratch: gradient of the largest activation
  float* dpool = nullptr;        // scratch: gradient of a pooled map
  float* wpart = nullptr;        // scratch: chunk partials of a conv's weight gradient
  float* loss = nullptr;
  int32_t* idx = nullptr;
  // dataset on the device
  float *d_masks = nullptr, *d_feat = nullptr, *d_targ = nullptr;
  int64_t d_count = 0;
  int64_t adam_t = 0;
};

namespace {

unsigned blocks_for(int64_t n) { return (unsigned)std::max<int64_t>(1, std::min<int64_t>((n + T - 1) / T, 65536)); }
// Workgroups of a matrix-core kernel for `tiles` 16 x 16 tiles (one wave each, no grid stride).
unsigned mma_blocks(int64_t tiles) { return (unsigned)std::max<int64_t>(1, (tiles + MW - 1) / MW); }
// Per conv layer and direction: 16-pixel tiles per wave and k pieces, so that a batch of 64
// samples gives each of the 1024 SIMDs a few waves.  The input gradient of layer 0 is unused.
struct ConvPlan {
  int pt, splits;
};
constexpr ConvPlan FWD_PLAN[NB] = {{4, 1}, {4, 1}, {1, 1}, {1, 4}};
constexpr ConvPlan BWD_PLAN[NB] = {{1, 1}, {4, 1}, {1, 2}, {1, 4}};
static_assert(SP[NB - 1] * SP[NB - 1] % 64 == 0, "every conv layer's side^2 must be a multiple of 16 * 4");

// k per piece of a layer's k range (a multiple of 4 * U).
int conv_kper(int K, int splits) {
  const int step = 4 * U;
  return ((K + splits - 1) / splits + step - 1) / step * step;
}

// Workgroups of T threads for one wave per item.
unsigned wave_blocks(int64_t items) { return (unsigned)std::max<int64_t>(1, (items + T / 64 - 1) / (T / 64)); }
// Batch-pixel terms per chunk of a conv's weight-gradient sum (a multiple of 4): 64 matrix-core
// steps per wave, and enough waves for the first layer (32 x 10 outputs).
constexpr int WCHUNK = 256;

// Weight-gradient column tiles per wave of layer l (the first layer has 10 columns, one tile),
// and the padded column count.
int conv_jt(int l) { return CH[l] * 9 + 1 > 16 ? 4 : 1; }
int conv_njp(int l) {
  const int w = 16 * conv_jt(l);
  return (CH[l] * 9 + 1 + w - 1) / w * w;
}

// Floats of the partial sums (weight-gradient chunks, conv k pieces) for batches of up to `b`
// samples.
int64_t wpart_floats(int64_t b) {
  int64_t m = 0;
  for (int l = 0; l < NB; ++l) {
    const int64_t njp = conv_njp(l);
    const int64_t nchunks = (b * SP[l] * SP[l] + WCHUNK - 1) / WCHUNK;
    m = std::max(m, nchunks * CH[l + 1] * njp);
    m = std::max(m, FWD_PLAN[l].splits * b * CH[l + 1] * SP[l] * SP[l]);
    m = std::max(m, BWD_PLAN[l].splits * b * CH[l] * SP[l] * SP[l]);
  }
  return m;
}

// One conv layer's forward pass (FWD) or input gradient through k_conv_mm (+ k_conv_split_sum).
template <bool FWD>
void launch_conv(vl_nn* q, int l, const float* src, float* dst, int n) {
  const ConvPlan pl = FWD ? FWD_PLAN[l] : BWD_PLAN[l];
  const int side = SP[l], S2 = side * side;
  const int M = FWD ? CH[l + 1] : CH[l], K = (FWD ? CH[l] : CH[l + 1]) * 9;
  const int kper = conv_kper(K, pl.splits);
  const int splits = (K + kper - 1) / kper;
  const int64_t waves = (int64_t)n * S2 / (16 * pl.pt) * (M / 16) * splits;
  const float* w = q->p + q->L.cw[l];
  const float* b = q->p + q->L.cb[l];
  if (pl.pt == 4)
    hipLaunchKernelGGL((k_conv_mm<4, FWD>), dim3(mma_blocks(waves)), dim3(MW * 64), 0, q->stream, src, w, b, dst,
                       q->wpart, n, CH[l], CH[l + 1], side, splits, kper);
  else
    hipLaunchKernelGGL((k_conv_mm<1, FWD>), dim3(mma_blocks(waves)), dim3(MW * 64), 0, q->stream, src, w, b, dst,
                       q->wpart, n, CH[l], CH[l + 1], side, splits, kper);
  if (splits > 1) {
    const int64_t total = (int64_t)n * M * S2;
    hipLaunchKernelGGL(k_conv_split_sum, dim3(blocks_for(total)), dim3(T), 0, q->stream, q->wpart, b, dst, total, M,
                       S2, splits, FWD ? 1 : 0);
  }
}

void free_nn(vl_nn* q) {
  if (!q) return;
  if (q->stream) (void)hipStreamSynchronize(q->stream);
  for (void* ptr : {(void*)q->p, (void*)q->g, (void*)q->m, (void*)q->v, (void*)q->x0, (void*)q->s, (void*)q->tgt,
                    (void*)q->h, (void*)q->y, (void*)q->dy, (void*)q->dh, (void*)q->dflat, (void*)q->dact,
                    (void*)q->dpool, (void*)q->wpart, (void*)q->loss, (void*)q->idx, (void*)q->d_masks, (void*)q->d_feat,
                    (void*)q->d_targ})
    if (ptr) (void)hipFree(ptr);
  for (int l = 0; l < NB; ++l) {
    if (q->act[l]) (void)hipFree(q->act[l]);
    if (q->pool[l]) (void)hipFree(q->pool[l]);
  }
  if (q->stream) (void)hipStreamDestroy(q->stream);
  delete q;
}

// forward pass over the gathered batch (x0, s) of n samples
vl_status forward(vl_nn* q, int n) {
  const float* in = q->x0;
  for (int l = 0; l < NB; ++l) {
    const int side = SP[l];
    launch_conv<true>(q, l, in, q->act[l], n);
    hipLaunchKernelGGL(k_pool_fwd, dim3(blocks_for((int64_t)n * CH[l + 1] * (side / 2) * (side / 2))), dim3(T), 0,
                       q->stream, q->act[l], q->pool[l], n * CH[l + 1], side);
    in = q->pool[l];
  }
  hipLaunchKernelGGL(k_fc_fwd, dim3(wave_blocks((int64_t)n * HID)), dim3(T), 0, q->stream, q->pool[NB - 1], q->s, FLAT,
                     q->p + q->L.w1, q->p + q->L.b1, q->h, n, HID, 1);
  hipLaunchKernelGGL(k_fc_fwd, dim3(wave_blocks((int64_t)n * OUT)), dim3(T), 0, q->stream, q->h, (const float*)nullptr,
                     HID, q->p + q->L.w2, q->p + q->L.b2, q->y, n, OUT, 0);
  VL_HIP(hipGetLastError());
  return VL_OK;
}

// backward pass into q->g (after forward and k_mse on the same batch)
vl_status backward(vl_nn* q, int n) {
  const Layout& L = q->L;
  hipLaunchKernelGGL(k_fc_bwd_w, dim3(blocks_for((int64_t)OUT * (HID + 1))), dim3(T), 0, q->stream, q->dy, q->h,
                     (const float*)nullptr, HID, q->g + L.w2, q->g + L.b2, n, OUT);
  hipLaunchKernelGGL(k_fc_bwd_x, dim3(blocks_for((int64_t)n * HID)), dim3(T), 0, q->stream, q->dy, q->p + L.w2, q->h,
                     HID, HID, q->dh, n, OUT, 1);
  hipLaunchKernelGGL(k_fc_bwd_w, dim3(blocks_for((int64_t)HID * (FIN + 1))), dim3(T), 0, q->stream, q->dh,
                     q->pool[NB - 1], q->s, FLAT, q->g + L.w1, q->g + L.b1, n, HID);
  hipLaunchKernelGGL(k_fc_bwd_x, dim3(blocks_for((int64_t)n * FLAT)), dim3(T), 0, q->stream, q->dh, q->p + L.w1,
                     q->pool[NB - 1], FIN, FLAT, q->dflat, n, HID, 0);
  // the pooled maps feed the next layer directly (their ReLU mask is applied in k_pool_bwd)
  const float* dp = q->dflat;
  for (int l = NB - 1; l >= 0; --l) {
    const int side = SP[l];
    hipLaunchKernelGGL(k_pool_bwd, dim3(blocks_for((int64_t)n * CH[l + 1] * side * side)), dim3(T), 0, q->stream,
                       q->act[l], dp, q->dact, n * CH[l + 1], side);
    const float* in = l == 0 ? q->x0 : q->pool[l - 1];
    const int njp = conv_njp(l), jt = conv_jt(l);
    const int nchunks = (int)(((int64_t)n * side * side + WCHUNK - 1) / WCHUNK);
    const unsigned wblocks = mma_blocks((int64_t)nchunks * (CH[l + 1] / 16) * (njp / (16 * jt)));
    if (jt == 4)
      hipLaunchKernelGGL(k_conv_bwd_w<4>, dim3(wblocks), dim3(MW * 64), 0, q->stream, q->dact, in, q->wpart, n, CH[l],
                         CH[l + 1], side, WCHUNK, nchunks, njp);
    else
      hipLaunchKernelGGL(k_conv_bwd_w<1>, dim3(wblocks), dim3(MW * 64), 0, q->stream, q->dact, in, q->wpart, n, CH[l],
                         CH[l + 1], side, WCHUNK, nchunks, njp);
    hipLaunchKernelGGL(k_conv_bwd_w_sum, dim3(wave_blocks((int64_t)CH[l + 1] * (CH[l] * 9 + 1))), dim3(T), 0,
                       q->stream, q->wpart, q->g + L.cw[l], q->g + L.cb[l], CH[l], CH[l + 1], nchunks, njp);
    if (l > 0) {
      launch_conv<false>(q, l, q->dact, q->dpool, n);
      dp = q->dpool;
    }
  }
  VL_HIP(hipGetLastError());
  return VL_OK;
}

vl_status gather_host(vl_nn* q, const float* masks, const float* feat, const float* targ, int n) {
  if (n < 1 || n > q->max_batch) return set_error(VL_EINVAL, "batch size must be 1..max_batch");
  VL_HIP(hipMemcpyAsync(q->x0, masks, (size_t)n * IMG * IMG * sizeof(float), hipMemcpyHostToDevice, q->stream));
  VL_HIP(hipMemcpyAsync(q->s, feat, (size_t)n * sizeof(float), hipMemcpyHostToDevice, q->stream));
  if (targ) VL_HIP(hipMemcpyAsync(q->tgt, targ, (size_t)n * OUT * sizeof(float), hipMemcpyHostToDevice, q->stream));
  return VL_OK;
}

}  // namespace

extern "C" {

vl_status vl_nn_create(int max_batch, uint64_t seed, vl_nn** out) {
  if (!out || max_batch < 1) return set_error(VL_EINVAL, "bad argument");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return set_error(VL_EDEVICE, "no HIP device");
  auto* q = new vl_nn;
  std::unique_ptr<vl_nn, void (*)(vl_nn*)> guard(q, free_nn);
  q->max_batch = max_batch;
  VL_HIP(hipStreamCreateWithFlags(&q->stream, hipStreamNonBlocking));
  const int64_t P = q->L.total;
  for (float** ptr : {&q->p, &q->g, &q->m, &q->v}) VL_HIP(hipMalloc((void**)ptr, (size_t)P * sizeof(float)));
  VL_HIP(hipMemset(q->g, 0, (size_t)P * sizeof(float)));
  VL_HIP(hipMemset(q->m, 0, (size_t)P * sizeof(float)));
  VL_HIP(hipMemset(q->v, 0, (size_t)P * sizeof(float)));
  const int64_t B = max_batch;
  VL_HIP(hipMalloc((void**)&q->x0, (size_t)(B * IMG * IMG) * sizeof(float)));
  VL_HIP(hipMalloc((void**)&q->s, (size_t)B * sizeof(float)));
  VL_HIP(hipMalloc((void**)&q->tgt, (size_t)(B * OUT) * sizeof(float)));
  int64_t biggest = 0;
  for (int l = 0; l < NB; ++l) {
    const int64_t a = B * CH[l + 1] * SP[l] * SP[l];
    biggest = std::max(biggest, a);
    VL_HIP(hipMalloc((void**)&q->act[l], (size_t)a * sizeof(float)));
    VL_HIP(hipMalloc((void**)&q->pool[l], (size_t)(a / 4) * sizeof(float)));
  }
  VL_HIP(hipMalloc((void**)&q->dact, (size_t)biggest * sizeof(float)));
  VL_HIP(hipMalloc((void**)&q->dpool, (size_t)biggest * sizeof(float)));
  VL_HIP(hipMalloc((void**)&q->wpart, (size_t)wpart_floats(B) * sizeof(float)));
  VL_HIP(hipMalloc((void**)&q->h, (size_t)(B * HID) * sizeof(float)));
  VL_HIP(hipMalloc((void**)&q->dh, (size_t)(B * HID) * sizeof(float)));
  VL_HIP(hipMalloc((void**)&q->y, (size_t)(B * OUT) * sizeof(float)));
  VL_HIP(hipMalloc((void**)&q->dy, (size_t)(B * OUT) * sizeof(float)));
  VL_HIP(hipMalloc((void**)&q->dflat, (size_t)(B * FLAT) * sizeof(float)));
  VL_HIP(hipMalloc((void**)&q->loss, sizeof(float)));
  VL_HIP(hipMalloc((void**)&q->idx, (size_t)B * sizeof(int32_t)));
  // He-normal weights (fan-in), zero biases, from the seed
  std::vector<float> hp((size_t)P, 0.0f);
  std::mt19937_64 rng(seed);
  auto fill = [&](int64_t off, int64_t count, int fan_in) {
    std::normal_distribution<double> nd(0.0, std::sqrt(2.0 / fan_in));
    for (int64_t k = 0; k < count; ++k) hp[(size_t)(off + k)] = (float)nd(rng);
  };
  for (int l = 0; l < NB; ++l) fill(q->L.cw[l], (int64_t)CH[l + 1] * CH[l] * 9, CH[l] * 9);
  fill(q->L.w1, (int64_t)HID * FIN, FIN);
  fill(q->L.w2, (int64_t)OUT * HID, HID);
  VL_HIP(hipMemcpy(q->p, hp.data(), (size_t)P * sizeof(float), hipMemcpyHostToDevice));
  *out = guard.release();
  return VL_OK;
}

void vl_nn_destroy(vl_nn* q) { free_nn(q); }

int64_t vl_nn_param_count(void) { return Layout().total; }

vl_status vl_nn_get_params(vl_nn* q, float* out) {
  if (!q || !out) return set_error(VL_EINVAL, "null argument");
  VL_HIP(hipStreamSynchronize(q->stream));
  VL_HIP(hipMemcpy(out, q->p, (size_t)q->L.total * sizeof(float), hipMemcpyDeviceToHost));
  return VL_OK;
}

vl_status vl_nn_set_params(vl_nn* q, const float* in) {
  if (!q || !in) return set_error(VL_EINVAL, "null argument");
  VL_HIP(hipStreamSynchronize(q->stream));
  VL_HIP(hipMemcpy(q->p, in, (size_t)q->L.total * sizeof(float), hipMemcpyHostToDevice));
  VL_HIP(hipMemset(q->m, 0, (size_t)q->L.total * sizeof(float)));
  VL_HIP(hipMemset(q->v, 0, (size_t)q->L.total * sizeof(float)));
  q->adam_t = 0;
  return VL_OK;
}

vl_status vl_nn_forward(vl_nn* q, const float* masks, const float* feat, int n, float* out) {
  if (!q || !masks || !feat || !out) return set_error(VL_EINVAL, "null argument");
  vl_status st = gather_host(q, masks, feat, nullptr, n);
  if (st != VL_OK) return st;
  st = forward(q, n);
  if (st != VL_OK) return st;
  VL_HIP(hipMemcpyAsync(out, q->y, (size_t)n * OUT * sizeof(float), hipMemcpyDeviceToHost, q->stream));
  VL_HIP(hipStreamSynchronize(q->stream));
  return VL_OK;
}

// Loss and parameter gradients of one batch (no update).
vl_status vl_nn_gradients(vl_nn* q, const float* masks, const float* feat, const float* targ, int n, float* loss,
                          float* grad) {
  if (!q || !masks || !feat || !targ || !loss || !grad) return set_error(VL_EINVAL, "null argument");
  vl_status st = gather_host(q, masks, feat, targ, n);
  if (st != VL_OK) return st;
  if ((st = forward(q, n)) != VL_OK) return st;
  hipLaunchKernelGGL(k_mse, dim3(1), dim3(T), 0, q->stream, q->y, q->tgt, q->dy, q->loss, n);
  if ((st = backward(q, n)) != VL_OK) return st;
  VL_HIP(hipMemcpyAsync(loss, q->loss, sizeof(float), hipMemcpyDeviceToHost, q->stream));
  VL_HIP(hipMemcpyAsync(grad, q->g, (size_t)q->L.total * sizeof(float), hipMemcpyDeviceToHost, q->stream));
  VL_HIP(hipStreamSynchronize(q->stream));
  return VL_OK;
}

// Keeps a dataset on the device: masks [count][64][64], feature [count], targets [count][3].
vl_status vl_nn_set_dataset(vl_nn* q, const float* masks, const float* feat, const float* targ, int64_t count) {
  if (!q || !masks || !feat || !targ || count < 1) return set_error(VL_EINVAL, "bad dataset");
  for (float* ptr : {q->d_masks, q->d_feat, q->d_targ})
    if (ptr) VL_HIP(hipFree(ptr));
  VL_HIP(hipMalloc((void**)&q->d_masks, (size_t)(count * IMG * IMG) * sizeof(float)));
  VL_HIP(hipMalloc((void**)&q->d_feat, (size_t)count * sizeof(float)));
  VL_HIP(hipMalloc((void**)&q->d_targ, (size_t)(count * OUT) * sizeof(float)));
  VL_HIP(hipMemcpy(q->d_masks, masks, (size_t)(count * IMG * IMG) * sizeof(float), hipMemcpyHostToDevice));
  VL_HIP(hipMemcpy(q->d_feat, feat, (size_t)count * sizeof(float), hipMemcpyHostToDevice));
  VL_HIP(hipMemcpy(q->d_targ, targ, (size_t)(count * OUT) * sizeof(float), hipMemcpyHostToDevice));
  q->d_count = count;
  return VL_OK;
}

// One pass over `count` dataset samples in the given order, in batches of `batch`, Adam
// (beta 0.9 / 0.999, eps 1e-8).  mean_loss = mean of the batch losses.
vl_status vl_nn_train_epoch(vl_nn* q, const int32_t* order, int64_t count, int batch, float lr, float* mean_loss) {
  if (!q || !order || !mean_loss || batch < 1 || batch > q->max_batch || count < 1)
    return set_error(VL_EINVAL, "bad argument");
  if (!q->d_masks) return set_error(VL_EINVAL, "no dataset on the device");
  for (int64_t k = 0; k < count; ++k)
    if (order[k] < 0 || order[k] >= q->d_count) return set_error(VL_EINVAL, "sample index out of range");
  int32_t* d_order = nullptr;
  VL_HIP(hipMalloc((void**)&d_order, (size_t)count * sizeof(int32_t)));
  std::unique_ptr<int32_t, void (*)(int32_t*)> og(d_order, [](int32_t* p) { (void)hipFree(p); });
  VL_HIP(hipMemcpyAsync(d_order, order, (size_t)count * sizeof(int32_t), hipMemcpyHostToDevice, q->stream));
  const int64_t nb = (count + batch - 1) / batch;
  float* d_losses = nullptr;
  VL_HIP(hipMalloc((void**)&d_losses, (size_t)nb * sizeof(float)));
  std::unique_ptr<float, void (*)(float*)> lg(d_losses, [](float* p) { (void)hipFree(p); });
  const Layout& L = q->L;
  for (int64_t b = 0; b < nb; ++b) {
    const int n = (int)std::min<int64_t>(batch, count - b * batch);
    hipLaunchKernelGGL(k_gather, dim3(blocks_for((int64_t)n * IMG * IMG)), dim3(T), 0, q->stream, q->d_masks, q->d_feat,
                       q->d_targ, d_order + b * batch, n, q->x0, q->s, q->tgt);
    vl_status st = forward(q, n);
    if (st != VL_OK) return st;
    hipLaunchKernelGGL(k_mse, dim3(1), dim3(T), 0, q->stream, q->y, q->tgt, q->dy, d_losses + b, n);
    if ((st = backward(q, n)) != VL_OK) return st;
    q->adam_t += 1;
    const float c1 = 1.0f - std::pow(0.9f, (float)q->adam_t), c2 = 1.0f - std::pow(0.999f, (float)q->adam_t);
    hipLaunchKernelGGL(k_adam, dim3(blocks_for(L.total)), dim3(T), 0, q->stream, q->p, q->g, q->m, q->v, L.total, lr,
                       0.9f, 0.999f, 1e-8f, c1, c2);
    VL_HIP(hipGetLastError());
  }
  std::vector<float> losses((size_t)nb);
  VL_HIP(hipMemcpyAsync(losses.data(), d_losses, (size_t)nb * sizeof(float), hipMemcpyDeviceToHost, q->stream));
  VL_HIP(hipStreamSynchronize(q->stream));
  double acc = 0.0;
  for (float v : losses) acc += v;
  *mean_loss = (float)(acc / (double)nb);
  return VL_OK;
}

// Predictions for dataset samples (by index) without copying masks again.
vl_status vl_nn_predict_dataset(vl_nn* q, const int32_t* order, int64_t count, float* out) {
  if (!q || !order || !out || count < 1) return set_error(VL_EINVAL, "bad argument");
  if (!q->d_masks) return set_error(VL_EINVAL, "no dataset on the device");
  for (int64_t k = 0; k < count; ++k)
    if (order[k] < 0 || order[k] >= q->d_count) return set_error(VL_EINVAL, "sample index out of range");
  for (int64_t b0 = 0; b0 < count; b0 += q->max_batch) {
    const int n = (int)std::min<int64_t>(q->max_batch, count - b0);
    VL_HIP(hipMemcpyAsync(q->idx, order + b0, (size_t)n * sizeof(int32_t), hipMemcpyHostToDevice, q->stream));
    hipLaunchKernelGGL(k_gather, dim3(blocks_for((int64_t)n * IMG * IMG)), dim3(T), 0, q->stream, q->d_masks, q->d_feat,
                       (const float*)nullptr, q->idx, n, q->x0, q->s, (float*)nullptr);
    vl_status st = forward(q, n);
    if (st != VL_OK) return st;
    VL_HIP(hipMemcpyAsync(out + b0 * OUT, q->y, (size_t)n * OUT * sizeof(float), hipMemcpyDeviceToHost, q->stream));
    VL_HIP(hipStreamSynchronize(q->stream));
  }
  return VL_OK;
}

}  // extern "C"
