// Host FP64 statistics of a force history: mean / rms, the Strouhal number from a padded
// Hann spectrum with log-parabolic peak refinement, and the mean-crossing frequency as an
// independent check (src/signal.rs:6-140; src/cases/channel.rs:182-206).  The estimators
// follow the reference's definitions so the reported Cd, Cl' and St mean the same thing.
#include <algorithm>
#include <cmath>
#include <complex>
#include <vector>

#include "setup.h"

namespace {

constexpr double kPi = 3.14159265358979323846;

// Iterative radix-2 decimation-in-time transform, forward sign e^{-i...}.
void fft_inplace(double* re, double* im, size_t n) {
  for (size_t i = 1, j = 0; i < n; ++i) {  // bit-reversal permutation
    size_t bit = n >> 1;
    for (; j & bit; bit >>= 1) j ^= bit;
    j |= bit;
    if (i < j) {
      std::swap(re[i], re[j]);
      std::swap(im[i], im[j]);
    }
  }
  for (size_t len = 2; len <= n; len <<= 1) {
    const double ang = -2.0 * kPi / (double)len;
    const double wr = std::cos(ang), wi = std::sin(ang);
    const size_t half = len / 2;
    for (size_t s = 0; s < n; s += len) {
      double cr = 1.0, ci = 0.0;  // twiddle by recurrence
      for (size_t k = 0; k < half; ++k) {
        const size_t a = s + k, b = a + half;
        const double tr = re[b] * cr - im[b] * ci, ti = re[b] * ci + im[b] * cr;
        re[b] = re[a] - tr;
        im[b] = im[a] - ti;
        re[a] += tr;
        im[a] += ti;
        const double ncr = cr * wr - ci * wi;
        ci = cr * wi + ci * wr;
        cr = ncr;
      }
    }
  }
}

double avg(const double* x, size_t n) {
  double s = 0.0;
  for (size_t i = 0; i < n; ++i) s += x[i];
  return s / (double)n;
}

double fluct_rms(const double* x, size_t n) {
  const double m = avg(x, n);
  double s = 0.0;
  for (size_t i = 0; i < n; ++i) s += (x[i] - m) * (x[i] - m);
  return std::sqrt(s / (double)n);
}

// (frequency, amplitude) of the strongest line of the Hann-windowed, mean-removed signal
// zero-padded to >= 8x its length; the peak bin is refined by a parabola through the log
// amplitudes of its neighbours.
void peak_frequency(const double* x, size_t n, double dt, double& f, double& amp) {
  const double m = avg(x, n);
  size_t size = 1;
  while (size < n * 8) size <<= 1;
  std::vector<double> re(size, 0.0), im(size, 0.0);
  for (size_t i = 0; i < n; ++i) re[i] = (x[i] - m) * (0.5 - 0.5 * std::cos(2.0 * kPi * (double)i / (double)n));
  fft_inplace(re.data(), im.data(), size);
  const double scale = 4.0 / (double)n;  // Hann coherent gain 1/2, one-sided
  const size_t bins = size / 2;
  auto amp_at = [&](size_t k) { return scale * std::sqrt(re[k] * re[k] + im[k] * im[k]); };
  auto freq_at = [&](size_t k) { return (double)k / ((double)size * dt); };
  size_t k = 1;
  for (size_t i = 1; i + 1 < bins; ++i)
    if (amp_at(i) > amp_at(k)) k = i;
  const double a0 = amp_at(k - 1), a1 = amp_at(k), a2 = k + 1 < bins ? amp_at(k + 1) : 0.0;
  if (k == 0 || k + 1 >= bins || a1 <= 0.0 || a0 <= 0.0 || a2 <= 0.0) {
    f = freq_at(k);
    amp = a1;
    return;
  }
  const double la = std::log(a0), lb = std::log(a1), lc = std::log(a2);
  const double shift = 0.5 * (la - lc) / (la - 2.0 * lb + lc);
  f = freq_at(k) + shift * (freq_at(1) - freq_at(0));
  amp = std::exp(lb - 0.25 * (la - lc) * shift);
}

// Mean spacing of upward crossings of the mean, with linear interpolation inside a sample.
bool upcross_frequency(const double* x, size_t n, double dt, double& f) {
  const double m = avg(x, n);
  size_t count = 0;
  double first = 0.0, last = 0.0;
  for (size_t i = 1; i < n; ++i) {
    const double a = x[i - 1] - m, b = x[i] - m;
    if (a < 0.0 && b >= 0.0) {
      const double t = (double)(i - 1) + a / (a - b);
      if (count == 0) first = t;
      last = t;
      ++count;
    }
  }
  if (count < 2) return false;
  f = (double)(count - 1) / ((last - first) * dt);
  return true;
}

}  // namespace

extern "C" {

vl_status vl_fft(double* re, double* im, size_t n) {
  if (!re || !im) return vl::set_error(VL_EINVAL, "null argument");
  if (n == 0 || (n & (n - 1))) return vl::set_error(VL_EINVAL, "FFT length must be a power of two");
  fft_inplace(re, im, n);
  return VL_OK;
}

vl_status vl_dominant_frequency(const double* x, size_t n, double dt, double* out2) {
  if (!x || !out2 || n < 2) return vl::set_error(VL_EINVAL, "need at least two samples");
  peak_frequency(x, n, dt, out2[0], out2[1]);
  return VL_OK;
}

// Local maxima above the mean, each refined by the parabola through it and its two
// neighbours: (fractional index, refined value).  All `count` peaks are counted; the first
// `cap` are written.
vl_status vl_peaks(const double* x, size_t n, double* pos, double* val, size_t cap, size_t* count) {
  if (!x || !count || (cap > 0 && (!pos || !val))) return vl::set_error(VL_EINVAL, "null argument");
  *count = 0;
  if (n < 3) return VL_OK;
  const double m = avg(x, n);
  size_t k = 0;
  for (size_t i = 1; i + 1 < n; ++i) {
    if (x[i] > m && x[i] > x[i - 1] && x[i] >= x[i + 1]) {
      const double a = x[i - 1], b = x[i], c = x[i + 1];
      const double denom = a - 2.0 * b + c;
      const double shift = denom != 0.0 ? 0.5 * (a - c) / denom : 0.0;
      if (k < cap) {
        pos[k] = (double)i + shift;
        val[k] = b - 0.25 * (a - c) * shift;
      }
      ++k;
    }
  }
  *count = k;
  return VL_OK;
}

vl_status vl_crossing_frequency(const double* x, size_t n, double dt, double* out, int* found) {
  if (!x || !out || !found) return vl::set_error(VL_EINVAL, "null argument");
  double f = 0.0;
  *found = upcross_frequency(x, n, dt, f) ? 1 : 0;
  *out = *found ? f : 0.0;
  return VL_OK;
}

// Series::stats: statistics of the samples from time `from` on; St only when the lift
// oscillates regularly (>= 3 periods in the window, FFT and crossing estimates within 5 %).
vl_status vl_series_stats(const double* cd_all, const double* cl_all, size_t len, double dt, double from,
                          vl_stats* out) {
  if (!cd_all || !cl_all || !out || len < 2) return vl::set_error(VL_EINVAL, "need at least two samples");
  const size_t i0 = std::min((size_t)(from / dt), len - 2);
  const double* cd = cd_all + i0;
  const double* cl = cl_all + i0;
  const size_t n = len - i0;
  out->cl_rms = fluct_rms(cl, n);
  out->st = 0.0;
  out->st_crossings = 0.0;
  out->periodic = 0;
  if (out->cl_rms > 1e-6) {
    double f, amp, fc = 0.0;
    peak_frequency(cl, n, dt, f, amp);
    if (!upcross_frequency(cl, n, dt, fc)) fc = 0.0;
    out->st_crossings = fc;
    const bool periodic = f * (double)n * dt >= 3.0 && std::fabs(fc / f - 1.0) < 0.05;
    out->periodic = periodic ? 1 : 0;
    if (periodic) out->st = f;
  }
  out->cd_mean = avg(cd, n);
  out->cd_rms = fluct_rms(cd, n);
  out->cl_mean = avg(cl, n);
  return VL_OK;
}

}  // extern "C"
