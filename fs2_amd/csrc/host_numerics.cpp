// FP64 host numerics: background cosmology, COLA time integrals, linear P(k).
//
// Replaces the GSL-based code of py/cosmology.cpp:24-171, py/cola.cpp:150-179
// and py/power.cpp:15-171.  Integrals use adaptive Gauss-Legendre quadrature
// whose nodes are computed here by Newton iteration on P_32 (no tabulated
// constants); the growth integral is taken in t = sqrt(a), which makes the
// integrand smooth at a = 0.  The table spline is a natural cubic spline of
// ln P in ln k, the interpolant GSL's gsl_interp_cspline builds.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <functional>
#include <string>
#include <vector>

#include "fs2_internal.h"

namespace fs2 {
namespace {

constexpr int kGL = 32;

struct GaussLegendre {
  double x[kGL], w[kGL];
  GaussLegendre() {
    for (int i = 0; i < kGL; ++i) {
      double z = std::cos(M_PI * (i + 0.75) / (kGL + 0.5));
      for (int it = 0; it < 100; ++it) {
        double p0 = 1.0, p1 = 0.0;
        for (int j = 1; j <= kGL; ++j) {
          const double p2 = p1;
          p1 = p0;
          p0 = ((2.0 * j - 1.0) * z * p1 - (j - 1.0) * p2) / j;
        }
        const double dp = kGL * (z * p0 - p1) / (z * z - 1.0);
        const double dz = p0 / dp;
        z -= dz;
        if (std::fabs(dz) < 1e-16) {
          x[i] = z;
          w[i] = 2.0 / ((1.0 - z * z) * dp * dp);
          break;
        }
        x[i] = z;
        w[i] = 2.0 / ((1.0 - z * z) * dp * dp);
      }
    }
  }
};

const GaussLegendre& gl() {
  static GaussLegendre g;
  return g;
}

double gl_rule(const std::function<double(double)>& f, double a, double b) {
  const GaussLegendre& g = gl();
  const double c = 0.5 * (a + b), h = 0.5 * (b - a);
  double s = 0.0;
  for (int i = 0; i < kGL; ++i) s += g.w[i] * f(c + h * g.x[i]);
  return s * h;
}

// Interval bisection until two levels agree to rtol relative or atol absolute;
// the depth bound keeps the cost finite when neither can be met.
double adapt(const std::function<double(double)>& f, double a, double b, double whole,
             double rtol, double atol, int depth) {
  const double m = 0.5 * (a + b);
  const double l = gl_rule(f, a, m), r = gl_rule(f, m, b);
  const double both = l + r;
  const double err = std::fabs(both - whole);
  if (depth <= 0 || err <= rtol * std::fabs(both) || err <= atol) return both;
  return adapt(f, a, m, l, rtol, 0.5 * atol, depth - 1) +
         adapt(f, m, b, r, rtol, 0.5 * atol, depth - 1);
}

double integrate(const std::function<double(double)>& f, double a, double b, double rtol = 1e-13,
                 double atol = 0.0) {
  if (a == b) return 0.0;
  const double whole = gl_rule(f, a, b);
  if (atol <= 0.0) atol = 1e-15 * std::fabs(whole);
  return adapt(f, a, b, whole, rtol, atol, 16);
}

// ---- cosmology state (py/cosmology.cpp:15-21) ----
double g_omega_m = 0.0;
double g_growth_norm = 0.0;

double growth_unnormalised(double a) {
  // cosmology.cpp:142-170: H(a) int_0^a (a H)^-3 da; with a = t^2 the
  // integrand becomes 2 t^4 (Om + (1-Om) t^6)^-3/2.
  const double om = g_omega_m;
  auto f = [om](double t) {
    const double t2 = t * t;
    const double u = om + (1.0 - om) * t2 * t2 * t2;
    return 2.0 * t2 * t2 / (u * std::sqrt(u));
  };
  return host::hubble(a) * integrate(f, 0.0, std::sqrt(a));
}

void check_init() {
  if (g_growth_norm == 0.0) fail(FS2_ERR_RUNTIME, "cosmology module not initialised");
}

}  // namespace

namespace host {

void cosmology_init(double omega_m) {
  if (!(omega_m > 0.0 && omega_m <= 1.0))
    fail(FS2_ERR_VALUE, "omega_m must be in (0, 1], got %g", omega_m);
  g_omega_m = omega_m;
  g_growth_norm = 1.0;
  g_growth_norm = 1.0 / growth_unnormalised(1.0);
}

bool cosmology_ready() { return g_growth_norm != 0.0; }

double omega_m() {
  check_init();
  return g_omega_m;
}

double hubble(double a) { return std::sqrt(g_omega_m / (a * a * a) + (1.0 - g_omega_m)); }

double D_growth(double a) {
  check_init();
  if (a == 0.0) return 0.0;
  return g_growth_norm * growth_unnormalised(a);
}

double omega(double a) {
  check_init();
  return g_omega_m / (g_omega_m + (1.0 - g_omega_m) * (a * a * a));
}

double D2_growth(double a, double D) {
  if (a == 0.0) return 0.0;
  return -3.0 / 7.0 * D * D * std::pow(omega(a), -1.0 / 143.0);
}

double f_growth_rate(double a) {
  check_init();
  if (a == 0.0) return 1.0;
  const double d_un = growth_unnormalised(a);
  const double hf = hubble(a);
  return 1.0 / (d_un * a * a * hf * hf) - 1.5 * g_omega_m / (hf * hf * a * a * a);
}

double Dv_growth(double a, double D) {
  if (a == 0.0) return 1.0;
  return a * a * D * hubble(a) * f_growth_rate(a);
}

double D2v_growth(double a, double D2) { return 2.0 * a * a * D2 * hubble(a) * f_growth_rate(a); }

// py/leapfrog.cpp:96-115: SphiStd = int_ai^af da / (a^2 H(a))
double leapfrog_Sphi(double ai, double af) {
  check_init();
  if (!(ai > 0.0)) fail(FS2_ERR_VALUE, "leapfrog kick from a = %g", ai);
  const double om = g_omega_m;
  auto f = [om](double a) { return 1.0 / (std::sqrt(om / (a * a * a) + 1.0 - om) * a * a); };
  return integrate(f, ai, af);
}

// py/leapfrog.cpp:117-141: SqStd = int_ai^af da / (a^3 H(a))
double leapfrog_Sq(double ai, double af) {
  check_init();
  if (!(ai > 0.0)) fail(FS2_ERR_VALUE, "leapfrog drift from a = %g", ai);
  const double om = g_omega_m;
  auto f = [om](double a) { return 1.0 / (std::sqrt(om / (a * a * a) + 1.0 - om) * a * a * a); };
  return integrate(f, ai, af);
}

// py/cola.cpp:152-179: int_ai^af a^n / (a^3 H(a)) da / av^n, n = -2.5
double drift_Sq(double ai, double af, double av) {
  check_init();
  if (!(ai > 0.0)) fail(FS2_ERR_VALUE, "drift from a = %g", ai);
  const double om = g_omega_m;
  const double n = -2.5;
  auto f = [om, n](double a) { return std::pow(a, n) / (std::sqrt(om / (a * a * a) + 1.0 - om) * a * a * a); };
  return integrate(f, ai, af) / std::pow(av, n);
}

}  // namespace host
}  // namespace fs2

// ---- power spectra -------------------------------------------------------------------
namespace fs2 {
namespace {

void spline_init(fs2_power* ps) {
  // natural cubic spline second derivatives (tridiagonal, m2[0] = m2[n-1] = 0)
  const int n = (int)ps->logk.size();
  const std::vector<double>& x = ps->logk;
  const std::vector<double>& y = ps->logP;
  ps->m2.assign(n, 0.0);
  if (n < 3) return;
  std::vector<double> c(n, 0.0), d(n, 0.0);
  for (int i = 1; i < n - 1; ++i) {
    const double h0 = x[i] - x[i - 1], h1 = x[i + 1] - x[i];
    const double a = h0, b = 2.0 * (h0 + h1), cc = h1;
    const double r = 6.0 * ((y[i + 1] - y[i]) / h1 - (y[i] - y[i - 1]) / h0);
    const double den = b - a * (i > 1 ? c[i - 1] : 0.0);
    c[i] = cc / den;
    d[i] = (r - a * (i > 1 ? d[i - 1] : 0.0)) / den;
  }
  for (int i = n - 2; i >= 1; --i) ps->m2[i] = d[i] - c[i] * ps->m2[i + 1];
}

double spline_eval(const fs2_power* ps, double lk) {
  const std::vector<double>& x = ps->logk;
  const int n = (int)x.size();
  int lo = 0, hi = n - 1;
  while (hi - lo > 1) {
    const int mid = (lo + hi) / 2;
    if (x[mid] > lk) hi = mid; else lo = mid;
  }
  const double h = x[hi] - x[lo];
  const double a = (x[hi] - lk) / h, b = (lk - x[lo]) / h;
  return a * ps->logP[lo] + b * ps->logP[hi] +
         ((a * a * a - a) * ps->m2[lo] + (b * b * b - b) * ps->m2[hi]) * h * h / 6.0;
}

double eh98_T(const fs2_power* ps, double k) {
  // Eisenstein & Hu 1998 eqs. 26-31 (no-wiggle), k in h/Mpc
  const double theta = ps->tcmb / 2.7;
  const double omh2 = ps->om * ps->h * ps->h, obh2 = ps->ob * ps->h * ps->h;
  const double fb = ps->ob / ps->om;
  const double s = 44.5 * std::log(9.83 / omh2) / std::sqrt(1.0 + 10.0 * std::pow(obh2, 0.75));
  const double alpha =
      1.0 - 0.328 * std::log(431.0 * omh2) * fb + 0.38 * std::log(22.3 * omh2) * fb * fb;
  const double ks = 0.43 * k * ps->h * s;
  const double gamma = ps->om * ps->h * (alpha + (1.0 - alpha) / (1.0 + ks * ks * ks * ks));
  const double q = k * theta * theta / gamma;
  const double L0 = std::log(2.0 * M_E + 1.8 * q);
  const double C0 = 14.2 + 731.0 / (1.0 + 62.5 * q);
  return L0 / (L0 + C0 * q * q);
}

double tophat_sigma2(const std::function<double(double)>& P, double R) {
  auto f = [&](double lnk) {
    const double k = std::exp(lnk);
    const double x = k * R;
    const double w = x > 1e-3 ? 3.0 * (std::sin(x) - x * std::cos(x)) / (x * x * x) : 1.0 - x * x / 10.0;
    return P(k) * w * w * k * k * k;
  };
  // segments of width 0.125 in ln k keep the top-hat oscillation resolved;
  // the absolute floor is relative to the whole integral
  const double lo = std::log(1e-6), hi = std::log(1e3);
  const int nseg = 160;
  std::vector<double> part(nseg);
  double rough = 0.0;
  for (int i = 0; i < nseg; ++i) {
    part[i] = gl_rule(f, lo + (hi - lo) * i / nseg, lo + (hi - lo) * (i + 1) / nseg);
    rough += std::fabs(part[i]);
  }
  double tot = 0.0;
  for (int i = 0; i < nseg; ++i)
    tot += integrate(f, lo + (hi - lo) * i / nseg, lo + (hi - lo) * (i + 1) / nseg, 1e-12,
                     1e-15 * rough);
  return tot / (2.0 * M_PI * M_PI);
}

}  // namespace

double power_P(const fs2_power* ps, double k) {
  if (ps->kind == 1) {
    const double T = eh98_T(ps, k);
    return ps->A * std::pow(k, ps->ns) * T * T;
  }
  const double lk = std::log(k);
  if (ps->logk.empty() || lk < ps->logk.front() || lk > ps->logk.back())
    fail(FS2_ERR_VALUE, "k = %g outside the power-spectrum table", k);
  return std::exp(spline_eval(ps, lk));
}

fs2_power* power_load(const char* filename) {
  // py/power.cpp:57-144
  FILE* fp = std::fopen(filename, "r");
  if (!fp) fail(FS2_ERR_IO, "unable to open power spectrum file: %s", filename);
  fs2_power* ps = new fs2_power();
  char line[128];
  double k_prev = 0.0;
  while (std::fgets(line, 127, fp)) {
    if (line[0] == '#') continue;
    double k, P;
    if (std::sscanf(line, "%lg %lg", &k, &P) != 2) break;
    if (k <= 0.0) continue;
    if (k < k_prev) {
      std::fclose(fp);
      delete ps;
      fail(FS2_ERR_IO, "k must be increasing in %s", filename);
    }
    ps->logk.push_back(std::log(k));
    ps->logP.push_back(std::log(P));
    k_prev = k;
  }
  std::fclose(fp);
  if (ps->logk.size() < 3) {
    delete ps;
    fail(FS2_ERR_IO, "not enough power spectrum data points in %s", filename);
  }
  spline_init(ps);
  return ps;
}

fs2_power* power_eh98(double om, double ob, double h, double ns, double sigma8) {
  if (!(om > 0 && ob >= 0 && ob < om && h > 0 && sigma8 > 0))
    fail(FS2_ERR_VALUE, "bad EH98 parameters");
  fs2_power* ps = new fs2_power();
  ps->kind = 1;
  ps->om = om; ps->ob = ob; ps->h = h; ps->ns = ns; ps->A = 1.0;
  const double s2 = tophat_sigma2([ps](double k) { return power_P(ps, k); }, 8.0);
  ps->A = sigma8 * sigma8 / s2;
  return ps;
}

double power_sigma(const fs2_power* ps, double R) {
  if (ps->kind == 1) return std::sqrt(tophat_sigma2([ps](double k) { return power_P(ps, k); }, R));
  // py/power.cpp:146-171 trapezoid over the table nodes
  const double fac = 1.0 / (2.0 * M_PI * M_PI);
  double k0 = std::exp(ps->logk[0]), f0 = 0.0, s2 = 0.0;
  for (size_t i = 0; i < ps->logk.size(); ++i) {
    const double k = std::exp(ps->logk[i]);
    const double x = k * R;
    const double w = 3.0 * (std::sin(x) - x * std::cos(x)) / (x * x * x);
    const double f1 = std::exp(ps->logP[i]) * k * k * w * w;
    s2 += 0.5 * (f0 + f1) * (k - k0);
    k0 = k;
    f0 = f1;
  }
  return std::sqrt(fac * s2);
}

}  // namespace fs2
