// Effective sample size by Geyer's initial monotone sequence estimator
// (C. J. Geyer, "Practical Markov chain Monte Carlo", Statist. Sci. 7, 1992):
// with autocorrelations rho_k of the trace, the sums of adjacent pairs
// Gamma_m = rho_2m + rho_2m+1 are positive, decreasing and convex for a
// reversible chain; the estimator sums them while positive, each clipped to
// the previous one, and tau = -1 + 2 sum_m Gamma_m, ESS = n / tau.
#include <cmath>
#include <vector>

#include "../../include/b3mcmc.h"

extern "C" double b3m_ess(const double *x, size_t n) {
  if (!x || n < 4) return 0.0;
  double mean = 0.0;
  for (size_t t = 0; t < n; t++) mean += x[t];
  mean /= (double)n;
  std::vector<double> d(n);
  for (size_t t = 0; t < n; t++) d[t] = x[t] - mean;
  auto autocov = [&](size_t k) {
    double s = 0.0;
    for (size_t t = 0; t + k < n; t++) s += d[t] * d[t + k];
    return s / (double)n;
  };
  const double g0 = autocov(0);
  if (!(g0 > 0.0)) return 0.0;
  double sum = 0.0, prev = INFINITY;
  for (size_t m = 0; 2 * m + 1 < n / 2; m++) {
    double gm = (autocov(2 * m) + autocov(2 * m + 1)) / g0;
    if (!(gm > 0.0)) break;
    if (gm > prev) gm = prev;  // monotone
    sum += gm;
    prev = gm;
  }
  const double tau = -1.0 + 2.0 * sum;
  if (!(tau > 0.0)) return (double)n;
  return (double)n / tau;
}
