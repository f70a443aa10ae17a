// Interpolation of time-varying inputs: piecewise constant, piecewise
// linear and natural cubic spline, for one series or several series that
// share their times.
//
// Replaces dust2::interpolate (dust2/interpolate/interpolate.hpp:14-269,
// dust2/interpolate/spline.hpp:10-90).  The knots (t, y) and the spline
// slopes k live in device memory; evaluation runs in the kernels
// (__host__ __device__), the slopes are solved once on the host.  Values
// are y[i * n_series + j] for knot i of series j (the layout of the
// reference's array interpolators); with one series the linear form is
// the reference's scalar one, ((y1 - y0) * (z - t0)) / (t1 - t0), with
// several its array one, y0 + (y1 - y0) * r, so that values agree bit for
// bit with the reference in both cases.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <stdexcept>
#include <string>
#include <vector>

namespace dust2_amd {

enum interp_kind : int { INTERP_CONSTANT = 0, INTERP_LINEAR = 1, INTERP_SPLINE = 2 };
enum interp_error : int { INTERP_OK = 0, INTERP_BEFORE = 1, INTERP_AFTER = 2 };

// Index i with t[i] <= z < t[i + 1]; z equal to the last time gives the
// last index, and so does z beyond it when the right-hand side may be
// extrapolated (piecewise constant only), otherwise INTERP_AFTER.  Times
// before the first (and NaN) give INTERP_BEFORE.
// dust2/interpolate/interpolate.hpp:17-46.
__host__ __device__ inline int interp_search(const double* t, int n, double z,
                                             bool extrapolate_rhs, int& err) {
  if (!(z >= t[0])) {
    err = INTERP_BEFORE;
    return 0;
  }
  int lo = 0, hi = n;  // first index with t[i] >= z
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (t[mid] < z) {
      lo = mid + 1;
    } else {
      hi = mid;
    }
  }
  if (lo == n) {
    if (!extrapolate_rhs) err = INTERP_AFTER;
    return n - 1;
  }
  return t[lo] != z ? lo - 1 : lo;
}

// Value of series j at time z.  k: spline slopes (same layout as y), used
// by INTERP_SPLINE only.
__host__ __device__ inline double interp_eval(int kind, const double* t, const double* y,
                                              const double* k, int n, int n_series, int j,
                                              double z, int& err) {
  const int i = interp_search(t, n, z, kind == INTERP_CONSTANT, err);
  if (err != INTERP_OK) return 0;
  const double yi = y[i * n_series + j];
  if (kind == INTERP_CONSTANT || i == n - 1) return yi;
  const double t0 = t[i], t1 = t[i + 1];
  const double y0 = yi, y1 = y[(i + 1) * n_series + j];
  if (kind == INTERP_LINEAR) {
    if (n_series == 1) return y0 + (y1 - y0) * (z - t0) / (t1 - t0);
    const double r = (z - t0) / (t1 - t0);
    return y0 + (y1 - y0) * r;
  }
  const double k0 = k[i * n_series + j], k1 = k[(i + 1) * n_series + j];
  const double r = (z - t0) / (t1 - t0);
  const double a = k0 * (t1 - t0) - (y1 - y0);
  const double b = -k1 * (t1 - t0) + (y1 - y0);
  return (1 - r) * y0 + r * y1 + r * (1 - r) * (a * (1 - r) + b * r);
}

// Knot validation with the reference's messages
// (dust2/interpolate/interpolate.hpp:48-66).
inline void interp_validate(const double* t, size_t n_t, size_t len_y, const char* name_t,
                            const char* name_y) {
  if (n_t != len_y) {
    throw std::runtime_error(std::string("Time variable '") + name_t +
                             "' and interpolation target '" + name_y +
                             "' must have the same length, but do not (" +
                             std::to_string(n_t) + " vs " + std::to_string(len_y) + ")");
  }
  for (size_t i = 1; i < n_t; ++i) {
    if (t[i - 1] >= t[i]) {
      throw std::runtime_error(std::string("Time variable '") + name_t +
                               "' must be strictly increasing but was not at index " +
                               std::to_string(i));
    }
  }
}

// Slopes k of the natural cubic spline through (t, y_j) for every series:
// the symmetric tridiagonal system of the first-derivative form, solved by
// forward elimination and back substitution
// (dust2/interpolate/spline.hpp:17-90).
inline std::vector<double> interp_spline_slopes(const double* t, const double* y, size_t n,
                                                size_t n_series) {
  if (n < 2) throw std::runtime_error("A spline needs at least two points");
  std::vector<double> lo(n), mid(n), up(n), rhs(n), k(n * n_series);
  const auto sq = [](double x) { return x * x; };
  mid[0] = 2 / (t[1] - t[0]);
  up[0] = 1 / (t[1] - t[0]);
  for (size_t i = 1; i + 1 < n; ++i) {
    lo[i] = 1 / (t[i] - t[i - 1]);
    mid[i] = 2 * (1 / (t[i] - t[i - 1]) + 1 / (t[i + 1] - t[i]));
    up[i] = 1 / (t[i + 1] - t[i]);
  }
  lo[n - 1] = 1 / (t[n - 1] - t[n - 2]);
  mid[n - 1] = 2 / (t[n - 1] - t[n - 2]);
  for (size_t j = 0; j < n_series; ++j) {
    const auto Y = [&](size_t i) { return y[i * n_series + j]; };
    rhs[0] = 3 * (Y(1) - Y(0)) / sq(t[1] - t[0]);
    for (size_t i = 1; i + 1 < n; ++i) {
      rhs[i] = 3 * ((Y(i) - Y(i - 1)) / sq(t[i] - t[i - 1]) +
                    (Y(i + 1) - Y(i)) / sq(t[i + 1] - t[i]));
    }
    rhs[n - 1] = 3 * (Y(n - 1) - Y(n - 2)) / sq(t[n - 1] - t[n - 2]);
    std::vector<double> diag(mid);
    for (size_t i = 1; i < n; ++i) {
      const double fac = lo[i] / diag[i - 1];
      diag[i] -= fac * up[i - 1];
      rhs[i] -= fac * rhs[i - 1];
    }
    rhs[n - 1] /= diag[n - 1];
    for (size_t i = n - 1; i-- > 0;) {
      rhs[i] = (rhs[i] - up[i] * rhs[i + 1]) / diag[i];
    }
    for (size_t i = 0; i < n; ++i) k[i * n_series + j] = rhs[i];
  }
  return k;
}

}  // namespace dust2_amd
