// Discrete-gamma rate categories (host, FP64), b3's update_categories
// (crates/b3/src/likelihood/gamma.rs:205-225): quantile midpoints of
// Gamma(alpha, 1), renormalised to mean 1.
//
// The quantile is found by Halley iteration on the regularised lower
// incomplete gamma P(a, x) (series for x < a + 1, Lentz continued fraction
// otherwise), started from the Wilson-Hilferty approximation, with
// bisection safeguards.
#include <cmath>
#include <string>

#include "../../include/b3hip.h"

namespace {

double lower_reg_gamma(double a, double x) {
  if (x <= 0.0) return 0.0;
  const double lg = std::lgamma(a);
  const double front = std::exp(a * std::log(x) - x - lg);
  if (x < a + 1.0) {
    double term = 1.0 / a, sum = term;
    for (int n = 1; n < 100000; n++) {
      term *= x / (a + n);
      sum += term;
      if (std::fabs(term) <= std::fabs(sum) * 1e-17) break;
    }
    return front * sum;
  }
  // continued fraction for Q(a, x)
  const double tiny = 1e-300;
  double b = x + 1.0 - a, cc = 1.0 / tiny, d = 1.0 / b, h = d;
  for (int i = 1; i < 100000; i++) {
    const double an = -i * (i - a);
    b += 2.0;
    d = an * d + b;
    if (std::fabs(d) < tiny) d = tiny;
    cc = b + an / cc;
    if (std::fabs(cc) < tiny) cc = tiny;
    d = 1.0 / d;
    const double del = d * cc;
    h *= del;
    if (std::fabs(del - 1.0) <= 1e-17) break;
  }
  return 1.0 - front * h;
}

double gamma_quantile(double a, double p) {
  // Wilson-Hilferty start: x ~ a (1 - t + z sqrt(t))^3, t = 1/(9a), with z
  // the standard normal quantile of p (found by bisection on erfc).
  const double t = 1.0 / (9.0 * a);
  double zq = 0.0;
  {
    double lo = -40.0, hi = 40.0;
    for (int i = 0; i < 200; i++) {
      const double mid = 0.5 * (lo + hi);
      const double cdf = 0.5 * std::erfc(-mid / std::sqrt(2.0));
      if (cdf < p) lo = mid; else hi = mid;
    }
    zq = 0.5 * (lo + hi);
  }
  double x = a * std::pow(1.0 - t + zq * std::sqrt(t), 3.0);
  if (!(x > 0.0) || !std::isfinite(x)) x = std::pow(p * std::tgamma(a + 1.0), 1.0 / a);
  if (!(x > 0.0) || !std::isfinite(x)) x = 1e-300;
  double lo = 0.0, hi = INFINITY;
  const double lg = std::lgamma(a);
  for (int it = 0; it < 400; it++) {
    const double f = lower_reg_gamma(a, x) - p;
    if (f < 0.0) lo = x; else hi = x;
    const double logdens = (a - 1.0) * std::log(x) - x - lg;
    const double dens = std::exp(logdens);
    double xn;
    if (dens > 0.0 && std::isfinite(dens)) {
      const double u = f / dens;
      const double corr = 1.0 - 0.5 * u * ((a - 1.0) / x - 1.0);  // Halley
      xn = x - u / (corr > 0.5 ? corr : 1.0);
    } else {
      xn = std::isfinite(hi) ? 0.5 * (lo + hi) : 2.0 * x;
    }
    if (!(xn > lo) || !(xn < hi)) xn = std::isfinite(hi) ? 0.5 * (lo + hi) : 2.0 * x;
    if (std::fabs(xn - x) <= 4e-16 * x) return xn;
    x = xn;
  }
  return x;
}

}  // namespace

extern "C" b3h_status b3h_gamma_rates(double alpha, uint32_t k, double *rates) {
  if (!rates || k == 0 || !(alpha > 0.0) || !std::isfinite(alpha)) return B3H_ERR_ARG;
  if (k == 1) {
    rates[0] = 1.0;
    return B3H_OK;
  }
  double sum = 0.0;
  for (uint32_t i = 0; i < k; i++) {
    const double p = (2.0 * i + 1.0) / (2.0 * k);
    rates[i] = gamma_quantile(alpha, p);
    sum += rates[i];
  }
  const double mean = sum / k;
  for (uint32_t i = 0; i < k; i++) rates[i] = rates[i] / mean;
  return B3H_OK;
}
