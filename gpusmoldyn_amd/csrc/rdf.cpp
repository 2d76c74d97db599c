// Radial distribution function (rdf) of a reactive pair under alternating
// diffusion and absorption steps, and the reaction-rate tables built from it
// (rxnparam.cpp:325-564; Andrews & Bray, Phys. Biol. 1:137, 2004).  FP64 on
// the host.  Lengths are reduced by the binding radius (absorbing sphere of
// radius 1); `step` is the reduced rms step sqrt(2 D dt) / a.
//
// One iteration: diffuse the rdf with the radial Green's function (the tail
// beyond the last point is taken as 1 + a2/r, a2 fitted to the outer 10% of
// the points), add the reverse-reaction source at the unbinding radius b when
// b >= 0, then absorb everything inside r = 1; the absorbed amount is the
// reaction flux per step.  Iterate to a steady flux.
#include <cmath>
#include <vector>

#include "smd_setup.h"

namespace smd {
namespace {

const double kPi = 3.14159265358979323846;
const double kSqrt2Pi = 2.50662827462;

// Integral of 4 pi r^2 f(r) over [r0, r1] for f linear between (r0, f0) and
// (r1, f1): the spherical trapezoid rule.
inline double shell(double r0, double r1, double f0, double f1) {
  return kPi * (f1 - f0) * (r1 + r0) * (r1 * r1 + r0 * r0) +
         4.0 * kPi / 3.0 * (r1 * f0 - r0 * f1) * (r1 * r1 + r1 * r0 + r0 * r0);
}

}  // namespace

// Mass inside the absorbing sphere (up to the first point past r = 1, minus
// the upper-left triangle of that last trapezoid), which is then removed.
double rdf_absorb(const double* r, double* rdf, int n) {
  double r0 = 0, r1 = 0, f0 = 0, f1 = rdf[0], mass = 0;
  int j = r[0] == 0 ? 1 : 0;
  for (; r1 < 1 && j < n; ++j) {
    r0 = r1;
    f0 = f1;
    r1 = r[j];
    f1 = rdf[j];
    mass += shell(r0, r1, f0, f1);
  }
  mass -= shell(r0, r1, 0.0, f1);
  for (int k = j - 2; k >= 0; --k) rdf[k] = 0;
  return mass;
}

// One diffusion step of length `step` applied to rdfa, written to rdfd.
void rdf_diffuse(const double* r, const double* rdfa, double* rdfd, int n, double step) {
  // tail 1 + a2 / r fitted over the outer 10% (least squares in 1/r)
  double alpha = 0, beta = 0;
  for (int i = (int)(0.9 * n); i < n; ++i) {
    alpha += 1.0 / (r[i] * r[i]);
    beta += (rdfa[i] - 1.0) / r[i];
  }
  const double a2 = beta / alpha / step;
  const double norm = 2.0 * kPi * kSqrt2Pi;
  int i = 0;
  if (r[0] == 0) {
    // at the origin the Green's function is a plain 3-D Gaussian
    double r0 = 0, r1 = 0, f0 = 0, f1 = 0, g = 0, acc = 0;
    for (int j = 1; j < n; ++j) {
      r0 = r1;
      f0 = f1;
      r1 = r[j] / step;
      g = std::exp(-r1 * r1 / 2.0) / norm;
      f1 = (rdfa[j] - rdfa[0]) * g;
      acc += shell(r0, r1, f0, f1);
    }
    acc += (1.0 - rdfa[0]) * (4.0 * kPi * r1 * g + erfc_d(r1 / std::sqrt(2.0)));
    rdfd[i++] = acc + rdfa[0];
  }
  for (; i < n; ++i) {
    const double x = r[i] / step;
    double r0 = 0, r1 = 0, f0 = 0, g = std::exp(-x * x / 2.0) / norm;
    double f1 = (rdfa[0] - rdfa[i]) * g, acc = 0;
    for (int j = r[0] == 0 ? 1 : 0; j < n; ++j) {
      r0 = r1;
      f0 = f1;
      r1 = r[j] / step;
      g = 1.0 / x / r1 * (std::exp(-(x - r1) * (x - r1) / 2.0) - std::exp(-(x + r1) * (x + r1) / 2.0)) /
          (2.0 * norm);
      f1 = (rdfa[j] - rdfa[i]) * g;
      acc += shell(r0, r1, f0, f1);
    }
    const double ed = erfc_d((r1 - x) / std::sqrt(2.0));
    const double es = erfc_d((r1 + x) / std::sqrt(2.0));
    acc += (1.0 - rdfa[i]) * (4.0 * kPi * r1 * g + 0.5 * (ed + es)) + a2 / 2.0 / x * (ed - es);
    rdfd[i] = acc + rdfa[i];
  }
}

// Adds the rdf of `flux` pairs released at separation b one diffusion step
// earlier (the analytic Green's function of a shell source).
void rdf_reverse(const double* r, double* rdf, int n, double step, double b, double flux) {
  const double k = 1.0 / (4.0 * kPi * kSqrt2Pi * step * step * step);
  if (b == 0) {
    for (int i = 0; i < n; ++i) {
      const double x = r[i] / step;
      rdf[i] += flux * k * 2.0 * std::exp(-x * x / 2.0);
    }
    return;
  }
  const double bs = b / step;
  int i = 0;
  if (r[0] == 0) rdf[i++] += flux * k * 2.0 * std::exp(-bs * bs / 2.0);
  for (; i < n; ++i) {
    const double x = r[i] / step;
    rdf[i] += flux * k / x / bs * (std::exp(-(x - bs) * (x - bs) / 2.0) - std::exp(-(x + bs) * (x + bs) / 2.0));
  }
}

double rdf_steady_state(const double* r, double* rdfa, double* rdfd, int n, double step, double b, double eps) {
  const int maxit = 100000;
  const double maxflux = 1e7;
  rdf_diffuse(r, rdfa, rdfd, n, step);
  double flux = rdf_absorb(r, rdfd, n), prev = flux;
  int it = 0;
  for (; it < 30 || (it < maxit && flux < maxflux && std::fabs((flux - prev) / (prev + 1e-20)) > eps); ++it) {
    prev = flux;
    rdf_diffuse(r, rdfa, rdfd, n, step);
    if (b >= 0) rdf_reverse(r, rdfd, n, step, b, flux);
    for (int i = 0; i < n; ++i) rdfa[i] = rdfd[i];
    flux = rdf_absorb(r, rdfa, n);
  }
  return (it >= maxit || flux >= maxflux) ? -1.0 : flux;
}

// Radial grid of rdfmaketable (rxnparam.cpp:500-505): spacing rmax / n from
// 0, the points next to 1 moved to 0.9999 and 1.0001.
void rdf_grid(double rmax, int n, double* r) {
  const double dr = rmax / n;
  r[0] = 0;
  int i = 1;
  for (; r[i - 1] < 1 && i - 1 < n; ++i) r[i] = r[i - 1] + dr;
  r[i - 1] = 0.9999;
  r[i++] = 1.0001;
  for (; i < n; ++i) r[i] = r[i - 1] + dr;
}

// One sweep of rdfmaketable over the 31 reduced steps exp(-3) exp(0.2)^k
// (increasing, or decreasing when `down`), each steady state starting from
// the previous one.  mode 'i': irreversible (b < 0, r up to 10); 'r': b > 1
// (r up to b + 3); 'b': 0 <= b <= 1 (r up to 5).
int rdf_sweep(char mode, int n, double eps, bool down, double b, double* flux31) {
  if (n < 10 || eps <= 0) return -1;
  const double rmax = mode == 'i' ? 10.0 : mode == 'r' ? b + 3.0 : 5.0;
  if (mode == 'i') b = -1.0;
  std::vector<double> r(n + 2), a(n + 2), d(n + 2);
  rdf_grid(rmax, n, r.data());
  int i = 0;
  for (; i < n && r[i] < 1; ++i) a[i] = 0;
  if (!down && mode == 'i')
    for (; i < n; ++i) a[i] = 1.0 - 1.0 / r[i];
  else if (!down && mode == 'r')
    for (; i < n && r[i] < b; ++i) a[i] = 1.0 - (b - r[i]) / r[i] / (b - 1.0);
  for (; i < n; ++i) a[i] = 1.0;
  // steps accumulate by repeated multiplication, as the table maker does
  const double inc = down ? 1.0 / std::exp(0.2) : std::exp(0.2);
  double s = down ? std::exp(3.0) : std::exp(-3.0), flux = 0;
  for (int k = 0; k < 31; ++k, s *= inc) {
    if (flux >= 0) flux = rdf_steady_state(r.data(), a.data(), d.data(), n, s, b, eps);
    flux31[down ? 30 - k : k] = flux;
  }
  return 0;
}

}  // namespace smd
