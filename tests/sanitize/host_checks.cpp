// Host-layer checks run under AddressSanitizer + UndefinedBehaviorSanitizer
// and under ThreadSanitizer (tests/test_host_sanitizers.py builds this file
// with g++ beside fs2_amd/csrc/fof.cpp and host_numerics.cpp, which need no
// GPU).  Exercises the multithreaded kd-tree build and union-find linking of
// the friends-of-friends finder (py/fof.cpp:45-161, py/kdtree.cpp:87-205)
// on clustered, degenerate and boundary inputs, its error path, and the
// FP64 cosmology, drift/kick integrals and power-spectrum loader
// (py/cosmology.cpp, py/cola.cpp:150-179, py/power.cpp:15-171) including the
// loader's IO errors.  Exits non-zero on a wrong result; the sanitizers
// abort on any memory error, undefined behaviour or data race.
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <random>
#include <string>
#include <vector>

#include "fs2_internal.h"

namespace fs2 {
void fail(int code, const char* fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  std::vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  throw Error{code, std::string(buf)};
}
fs2_power* power_load(const char* filename);
fs2_power* power_eh98(double om, double ob, double h, double ns, double sigma8);
double power_sigma(const fs2_power* ps, double R);
}  // namespace fs2

static int g_failures = 0;
#define CHECK(cond)                                                   \
  do {                                                                \
    if (!(cond)) {                                                    \
      std::fprintf(stderr, "CHECK failed: %s (line %d)\n", #cond, __LINE__); \
      ++g_failures;                                                   \
    }                                                                 \
  } while (0)

template <typename F>
static int error_code(F f) {
  try {
    f();
  } catch (const fs2::Error& e) {
    return e.code;
  }
  return 0;
}

// Labels must be the lowest index of each linked set: grp[i] <= i,
// grp[grp[i]] == grp[i], and every pair closer than ll (periodic, FP32
// distance as in fof.cpp:21-37) shares a label -- brute force over all pairs.
static void check_fof(const std::vector<float>& x, double L, double ll, int quota,
                      const double* box3) {
  const int64_t n = (int64_t)x.size() / 3;
  std::vector<float> planes(3 * n);
  for (int64_t i = 0; i < n; ++i)
    for (int a = 0; a < 3; ++a) planes[a * n + i] = x[3 * i + a];
  const float* xyz[3] = {planes.data(), planes.data() + n, planes.data() + 2 * n};
  fs2::FofResult r;
  fs2::fof_groups(xyz, n, n, L, ll, box3, quota, r);
  CHECK((int64_t)r.grp.size() == n && (int64_t)r.perm.size() == n);
  std::vector<int> seen(n, 0);
  for (int64_t i = 0; i < n; ++i) {
    CHECK(r.perm[i] >= 0 && r.perm[i] < n);
    if (r.perm[i] >= 0 && r.perm[i] < n) seen[r.perm[i]]++;
    CHECK(r.grp[i] >= 0 && r.grp[i] <= i && r.grp[r.grp[i]] == r.grp[i]);
  }
  for (int64_t i = 0; i < n; ++i) CHECK(seen[i] == 1);
  const float Lf = (float)L, hL = (float)(L / 2), ll2 = (float)(ll * ll);
  for (int64_t i = 0; i < n; ++i) {
    const float* a = &x[3 * r.perm[i]];
    for (int64_t j = i + 1; j < n; ++j) {
      const float* b = &x[3 * r.perm[j]];
      float d2 = 0.f;
      for (int k = 0; k < 3; ++k) {
        float d = std::fabs(a[k] - b[k]);
        if (d > hL) d = Lf - d;
        d2 += d * d;
      }
      if (d2 < ll2) CHECK(r.grp[i] == r.grp[j]);
    }
  }
  std::vector<int32_t> nfof;
  fs2::fof_counts(r.grp, nfof);
  int64_t tot = 0;
  for (int32_t c : nfof) tot += c;
  CHECK(tot == n);
}

int main(int argc, char** argv) {
  const char* table = argc > 1 ? argv[1] : nullptr;
  std::mt19937_64 rng(12345);
  std::uniform_real_distribution<double> U(0.0, 1.0);
  std::normal_distribution<double> G(0.0, 1.0);

  // clustered particles in a periodic box, three quotas, with and without
  // the caller's box extents
  {
    const double L = 50.0;
    std::vector<float> x;
    for (int c = 0; c < 40; ++c) {
      const double cx = U(rng) * L, cy = U(rng) * L, cz = U(rng) * L;
      for (int k = 0; k < 60; ++k) {
        const double p[3] = {cx + 0.8 * G(rng), cy + 0.8 * G(rng), cz + 0.8 * G(rng)};
        for (double q : p) x.push_back((float)std::fmod(std::fmod(q, L) + L, L));
      }
    }
    for (int k = 0; k < 600; ++k)
      for (int a = 0; a < 3; ++a) x.push_back((float)(U(rng) * L));
    const double box3[3] = {L, L, L};
    check_fof(x, L, 0.4, 1, nullptr);
    check_fof(x, L, 0.4, 8, box3);
    check_fof(x, L, 0.9, 100000, nullptr);
  }
  // degenerate sets: none, one, all particles at one point, on the faces
  {
    check_fof({}, 10.0, 1.0, 4, nullptr);
    check_fof({1.f, 2.f, 3.f}, 10.0, 1.0, 4, nullptr);
    std::vector<float> same(3 * 500, 2.5f);
    check_fof(same, 10.0, 0.1, 2, nullptr);
    std::vector<float> faces;
    for (int k = 0; k < 300; ++k) {
      const float e = (k & 1) ? 0.0f : std::nextafter(10.0f, 0.0f);
      faces.push_back(e);
      faces.push_back((float)(U(rng) * 10.0));
      faces.push_back((k & 2) ? 0.0f : (float)(U(rng) * 10.0));
    }
    check_fof(faces, 10.0, 0.3, 3, nullptr);
  }
  // error path: quota < 1
  {
    std::vector<float> x(30, 1.0f);
    const float* xyz[3] = {x.data(), x.data() + 10, x.data() + 20};
    fs2::FofResult r;
    CHECK(error_code([&] { fs2::fof_groups(xyz, 10, 10, 5.0, 1.0, nullptr, 0, r); }) ==
          FS2_ERR_VALUE);
  }

  // cosmology and time-step integrals
  CHECK(error_code([] { fs2::host::D_growth(0.5); }) == FS2_ERR_RUNTIME);
  CHECK(error_code([] { fs2::host::cosmology_init(1.5); }) == FS2_ERR_VALUE);
  fs2::host::cosmology_init(0.31);
  CHECK(std::fabs(fs2::host::D_growth(1.0) - 1.0) < 1e-12);
  CHECK(std::fabs(fs2::host::D_growth(0.1) - 0.127351097784) < 1e-9);
  for (double a = 0.05; a <= 1.0; a += 0.05) {
    const double D = fs2::host::D_growth(a);
    CHECK(D > 0 && std::isfinite(fs2::host::D2_growth(a, D)));
    CHECK(std::isfinite(fs2::host::Dv_growth(a, D)) && fs2::host::f_growth_rate(a) > 0);
    CHECK(fs2::host::drift_Sq(a, a + 0.05, a + 0.02) > 0);
    CHECK(fs2::host::leapfrog_Sphi(a, a + 0.05) > 0 && fs2::host::leapfrog_Sq(a, a + 0.05) > 0);
  }
  CHECK(error_code([] { fs2::host::drift_Sq(0.0, 0.1, 0.1); }) == FS2_ERR_VALUE);

  // power spectra: the reference's table, IO errors, EH98
  if (table) {
    fs2_power* ps = fs2::power_load(table);
    CHECK(ps->logk.size() > 100);
    const double s8 = fs2::power_sigma(ps, 8.0);
    CHECK(s8 > 0.7 && s8 < 0.9);
    CHECK(fs2::power_P(ps, 0.1) > 0);
    CHECK(error_code([&] { fs2::power_P(ps, 1e9); }) == FS2_ERR_VALUE);
    delete ps;
  }
  CHECK(error_code([] { fs2::power_load("/nonexistent/pk.dat"); }) == FS2_ERR_IO);
  {
    const char* bad = "host_checks_bad_pk.dat";
    FILE* f = std::fopen(bad, "w");
    std::fprintf(f, "# k P\n0.1 10\n0.05 20\n0.2 5\n");
    std::fclose(f);
    CHECK(error_code([&] { fs2::power_load(bad); }) == FS2_ERR_IO);
    f = std::fopen(bad, "w");
    std::fprintf(f, "0.1 10\nnot a number\n");
    std::fclose(f);
    CHECK(error_code([&] { delete fs2::power_load(bad); }) == FS2_ERR_IO);
    std::remove(bad);
  }
  {
    fs2_power* eh = fs2::power_eh98(0.31, 0.049, 0.67, 0.96, 0.81);
    CHECK(std::fabs(fs2::power_sigma(eh, 8.0) - 0.81) < 1e-6);
    delete eh;
    CHECK(error_code([] { fs2::power_eh98(-1, 0.049, 0.67, 0.96, 0.81); }) == FS2_ERR_VALUE);
  }
  if (g_failures) std::fprintf(stderr, "%d checks failed\n", g_failures);
  else std::printf("host checks passed\n");
  return g_failures ? 1 : 0;
}
