// C ABI of libfs2hip: handles, error codes, device buffers and the host-side
// orchestration of the PM / COLA / 2LPT hot path.  See include/fs2hip.h.
#include <algorithm>
#include <memory>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "fs2_internal.h"

// ---- handles ----------------------------------------------------------------------------
struct fs2_grid {
  int nc = 0;
  float* d = nullptr;   // padded [nc][nc][2(nc/2+1)]
  bool owning = true;
  fs2_pm* owner = nullptr;   // the PM whose density mesh this view shows
};

struct fs2_particles {
  uint64_t np_alloc = 0, np_local = 0, np_total = 0;
  double boxsize = 0.0;
  double a_x = 0.0, a_v = 0.0, a_f = 0.0;
  float* planes = nullptr;   // x, v, dx1, dx2, force: each [3][np_alloc]
  uint64_t* id = nullptr;
  bool wrap_pending = false;
  uint64_t x_version = 0;    // changes whenever positions change (see touch_x)
  // A COLA kick waiting to run with the next drift as one fused pass; any
  // other access to the particles applies it first (apply_kick).  Once the
  // device planes have been handed out, kicks run at once.
  struct {
    bool pending = false;
    double m15om = 0.0;
    float kf = 0.f, q1 = 0.f, q2 = 0.f;
  } kick;
  // A COLA drift (with the kick recorded before it, if any) waiting for the
  // next density binning, which runs it in the same pass over the particles
  // (paint_tiled_bin_kick_drift); any other access applies it first.  Once
  // the device planes have been handed out, drifts run at once.
  struct {
    bool pending = false;
    double dt = 0.0;
    float da1 = 0.f, da2 = 0.f;
  } drift;
  bool planes_exposed = false;
  float* field(int k) const { return planes + (size_t)k * 3 * np_alloc; }
  float* x() const { return field(0); }
  float* v() const { return field(1); }
  float* dx1() const { return field(2); }
  float* dx2() const { return field(3); }
  float* f() const { return field(4); }
};

enum PmStatus { PM_DONE = 0, PM_DENSITY_DONE = 1, PM_FORCE_DONE = 2 };   // py/pm.h:8

struct fs2_pm {
  int nc = 0;
  double pm_factor = 1.0;
  double boxsize = 0.0;
  float boxsize_f = 0.f;
  float* mesh = nullptr;      // density, then delta_k in place
  float* force[3] = {nullptr, nullptr, nullptr};
  fs2_grid density;           // borrowed view of mesh
  int status = PM_DONE;
  int variant = FS2_PAINT_SORTED;
  bool mesh_is_k = false;
  fs2::PaintScratch paint;     // tiled-paint scratch, grown on demand
  fs2::TileOpts tiles;         // paint tiles: column tiles where the fused readout runs
  uint64_t binned_version = 0; // x_version of the particles the scratch bins
  uint64_t binned_np = 0;
  // force[a] hold the spectra before their z c2r (fused readout); the real
  // meshes are made on demand
  bool force_spectral = false;
  // compute_density binned the particles but left the tile kernel for
  // compute_force to fuse with the z r2c; any other use of the density mesh
  // runs it first (pm_materialize)
  bool density_pending = false;
};

namespace fs2 {
fs2_power* power_load(const char* filename);
fs2_power* power_eh98(double om, double ob, double h, double ns, double sigma8);
double power_sigma(const fs2_power* ps, double R);

// ---- errors -----------------------------------------------------------------------------
namespace {
thread_local std::string g_err;
}

void set_error(int code, const std::string& msg) {
  (void)code;
  g_err = msg;
}

void fail(int code, const char* fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  std::vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  throw Error{code, std::string(buf)};
}


// ---- runtime ----------------------------------------------------------------------------
namespace {
std::mutex g_init_mu;
bool g_inited = false;
hipStream_t g_stream = nullptr;
int g_loglevel = 2;
}

void device_init() {
  std::lock_guard<std::mutex> lock(g_init_mu);
  if (g_inited) return;
  int n = 0;
  hipError_t e = hipGetDeviceCount(&n);
  if (e != hipSuccess || n == 0)
    fail(FS2_ERR_DEVICE, "no HIP device available (hipGetDeviceCount: %s, n=%d)",
         hipGetErrorString(e), n);
  FS2_HIP(hipSetDevice(0));
  FS2_HIP(hipStreamCreateWithFlags(&g_stream, hipStreamNonBlocking));
  pk_scratch_init();   // device 0 is current
  g_inited = true;
}

namespace {
thread_local hipStream_t t_stream_override = nullptr;
}

hipStream_t stream() {
  if (t_stream_override) return t_stream_override;
  if (!g_inited) device_init();
  return g_stream;
}

int device_cus() {
  static int cus = 0;
  if (cus == 0) {
    device_init();
    int v = 0;
    FS2_HIP(hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, 0));
    cus = v > 0 ? v : 1;
  }
  return cus;
}

hipStream_t library_stream() {
  if (!g_inited) device_init();
  return g_stream;
}

StreamScope::StreamScope(hipStream_t s) : prev_(t_stream_override) { t_stream_override = s; }
StreamScope::~StreamScope() { t_stream_override = prev_; }

void* dmalloc(size_t bytes) {
  device_init();
  void* p = nullptr;
  if (bytes == 0) bytes = 256;
  hipError_t e = hipMalloc(&p, bytes);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    fail(FS2_ERR_MEMORY, "device allocation of %zu bytes failed: %s", bytes, hipGetErrorString(e));
  }
  return p;
}

void dfree(void* p) {
  if (p) (void)hipFree(p);
}

void sync() { FS2_HIP(hipStreamSynchronize(stream())); }

int loglevel() { return g_loglevel; }

void logf(int level, const char* fmt, ...) {
  if (level < g_loglevel) return;
  va_list ap;
  va_start(ap, fmt);
  std::vfprintf(stderr, fmt, ap);
  va_end(ap);
}

// ---- kernel timing ----------------------------------------------------------------------------
namespace {
struct ProfRec {
  std::string name;
  uint64_t launches = 0;
  double total_ms = 0.0;
  double bytes = 0.0;
};
struct Pending {
  int rec;
  hipEvent_t a, b;
};
bool g_prof = false;
std::vector<ProfRec> g_recs;
std::vector<Pending> g_pending;
std::vector<hipEvent_t> g_event_pool;

hipEvent_t get_event() {
  if (!g_event_pool.empty()) {
    hipEvent_t e = g_event_pool.back();
    g_event_pool.pop_back();
    return e;
  }
  hipEvent_t e;
  FS2_HIP(hipEventCreate(&e));
  return e;
}

void prof_resolve() {
  if (g_pending.empty()) return;
  FS2_HIP(hipStreamSynchronize(stream()));
  for (const Pending& p : g_pending) {
    float ms = 0.f;
    FS2_HIP(hipEventElapsedTime(&ms, p.a, p.b));
    g_recs[p.rec].total_ms += ms;
    g_recs[p.rec].launches += 1;
    g_event_pool.push_back(p.a);
    g_event_pool.push_back(p.b);
  }
  g_pending.clear();
}
}  // namespace

int prof_begin(const char* name, double bytes) {
  if (!g_prof) return -1;
  int rec = -1;
  for (size_t i = 0; i < g_recs.size(); ++i)
    if (g_recs[i].name == name) rec = (int)i;
  if (rec < 0) {
    g_recs.push_back(ProfRec{name, 0, 0.0, bytes});
    rec = (int)g_recs.size() - 1;
  }
  g_recs[rec].bytes = bytes;
  Pending p{rec, get_event(), get_event()};
  FS2_HIP(hipEventRecord(p.a, stream()));
  g_pending.push_back(p);
  if (g_pending.size() > 4096) prof_resolve();
  return (int)g_pending.size() - 1;
}

void prof_end(int token) {
  if (token < 0 || !g_prof) return;
  FS2_HIP(hipEventRecord(g_pending[token].b, stream()));
}

bool profiling_enabled() { return g_prof; }
void set_profiling(bool on) { g_prof = on; }

// ---- helpers ----------------------------------------------------------------------------------
namespace {


void apply_kick(fs2_particles* p) {
  if (!p->kick.pending) return;
  cola_kick(p->v(), p->f(), p->dx1(), p->dx2(), (int64_t)p->np_alloc, (int64_t)p->np_local,
            p->kick.m15om, p->kick.kf, p->kick.q1, p->kick.q2);
  p->kick.pending = false;   // only once launched: a failed launch leaves it owed
}

// The pending drift and the kick recorded before it as one update.
KickDrift pending_kick_drift(const fs2_particles* p) {
  KickDrift c{};
  c.kick = p->kick.pending ? 1 : 0;
  c.m15om = p->kick.m15om;
  c.kf = p->kick.kf;
  c.q1 = p->kick.q1;
  c.q2 = p->kick.q2;
  c.dt = p->drift.dt;
  c.da1 = p->drift.da1;
  c.da2 = p->drift.da2;
  return c;
}

// Called once the pending update has been launched.
void clear_pending(fs2_particles* p) {
  p->kick.pending = false;
  p->drift.pending = false;
}

// Runs whatever COLA update is pending: kick and drift in one pass, the
// drift alone, or the kick alone.
void apply_pending(fs2_particles* p) {
  if (!p->drift.pending) {
    apply_kick(p);
    return;
  }
  const KickDrift c = pending_kick_drift(p);
  const int64_t s = (int64_t)p->np_alloc, n = (int64_t)p->np_local;
  if (c.kick)
    cola_kick_drift(p->x(), p->x(), p->v(), p->f(), p->dx1(), p->dx2(), s, n, c.m15om, c.kf, c.q1,
                    c.q2, c.dt, c.da1, c.da2);
  else
    cola_drift(p->x(), p->v(), p->dx1(), p->dx2(), s, n, c.dt, c.da1, c.da2);
  clear_pending(p);
}

// Validates the handle and brings the device planes up to date: every entry
// point that reads or writes particle data goes through here.
fs2_particles* check_p(fs2_particles* p) {
  if (!p) fail(FS2_ERR_VALUE, "null particles handle");
  apply_pending(p);
  return p;
}

PosArgs pos_of(fs2_particles* p) { return PosArgs{p->x(), (int64_t)p->np_alloc, (int64_t)p->np_local}; }

// Positions changed: give them a version no other state ever had, so bins
// computed from earlier positions are never reused.
uint64_t g_version_counter = 0;
void touch_x(fs2_particles* p) { p->x_version = ++g_version_counter; }

void flush_wrap(fs2_particles* p) {
  if (p->wrap_pending) {
    wrap_positions(pos_of(p), (float)p->boxsize);
    p->wrap_pending = false;
    touch_x(p);
  }
}

// Device planes [3][stride] -> host rows (n, 3)
void download_rows(const float* d, uint64_t stride, uint64_t n, float* out) {
  if (n == 0) return;
  std::vector<float> tmp(3 * n);
  for (int a = 0; a < 3; ++a)
    FS2_HIP(hipMemcpyAsync(tmp.data() + a * n, d + a * stride, n * sizeof(float),
                           hipMemcpyDeviceToHost, stream()));
  sync();
  for (uint64_t i = 0; i < n; ++i)
    for (int a = 0; a < 3; ++a) out[3 * i + a] = tmp[a * n + i];
}

void upload_rows(float* d, uint64_t stride, uint64_t n, const float* in) {
  if (n == 0) return;
  std::vector<float> tmp(3 * n);
  for (uint64_t i = 0; i < n; ++i)
    for (int a = 0; a < 3; ++a) tmp[a * n + i] = in[3 * i + a];
  for (int a = 0; a < 3; ++a)
    FS2_HIP(hipMemcpyAsync(d + a * stride, tmp.data() + a * n, n * sizeof(float),
                           hipMemcpyHostToDevice, stream()));
  sync();
}

fs2_grid* grid_new(int nc) {
  if (nc < 2 || !fft_supported(nc)) fail(FS2_ERR_VALUE, "grid size %d not supported", nc);
  fs2_grid* g = new fs2_grid();
  g->nc = nc;
  g->d = (float*)dmalloc(grid_floats(nc) * sizeof(float));
  FS2_HIP(hipMemsetAsync(g->d, 0, grid_floats(nc) * sizeof(float), stream()));
  return g;
}

void grid_get_real(const fs2_grid* g, float* out) {
  const int n = g->nc, ncz = grid_rpitch(n);
  FS2_HIP(hipMemcpy2DAsync(out, n * sizeof(float), g->d, ncz * sizeof(float), n * sizeof(float),
                           (size_t)n * n, hipMemcpyDeviceToHost, stream()));
  sync();
}

void grid_set_real(fs2_grid* g, const float* in) {
  const int n = g->nc, ncz = grid_rpitch(n);
  FS2_HIP(hipMemcpy2DAsync(g->d, ncz * sizeof(float), in, n * sizeof(float), n * sizeof(float),
                           (size_t)n * n, hipMemcpyHostToDevice, stream()));
  sync();
}

float lpt_offset = 0.5f;   // py/lpt.cpp:29

}  // namespace

void free_paint_scratch(PaintScratch& s) {
  dfree(s.counts);
  dfree(s.offsets);
  dfree(s.cursor);
  dfree(s.block_sums);
  dfree(s.rec);
  dfree(s.amax);
  s = PaintScratch();
}

void ensure_paint_scratch(PaintScratch& s, size_t np, size_t nbins) {
  if (s.rec && s.np >= np && s.nbins == nbins) return;
  if (g_inited) FS2_HIP(hipStreamSynchronize(g_stream));
  free_paint_scratch(s);
  const size_t cap = std::max<size_t>(np, 1);
  s.counts = (unsigned*)dmalloc(nbins * sizeof(unsigned));
  s.offsets = (unsigned*)dmalloc((nbins + 1) * sizeof(unsigned));
  s.cursor = (unsigned*)dmalloc(nbins * sizeof(unsigned));
  s.block_sums = (unsigned*)dmalloc(4096 * sizeof(unsigned));
  s.rec = (float4*)dmalloc(cap * sizeof(float4));
  s.amax = (unsigned*)dmalloc(4 * sizeof(unsigned));
  s.np = cap;
  s.nbins = nbins;
}

float lpt_offset_value() { return lpt_offset; }

std::vector<double> lpt_amp_table(int n, double boxsize, const fs2_power* ps) {
  const int h = n / 2;
  const int nk2 = 3 * h * h + 1;
  std::vector<double> amp(nk2, 0.0);
  const double kf = 2.0 * M_PI / boxsize;
  const double V = boxsize * boxsize * boxsize;
  const double norm = 1.0 / std::pow((double)n, 1.5);
  std::vector<char> used(nk2, 0);
  for (int ix = -h + 1; ix < h; ++ix)
    for (int iy = -h + 1; iy < h; ++iy)
      for (int iz = 0; iz < h; ++iz) used[ix * ix + iy * iy + iz * iz] = 1;
  for (int q = 1; q < nk2; ++q)
    if (used[q]) amp[q] = std::sqrt(power_P(ps, kf * std::sqrt((double)q)) / V) * norm;
  return amp;
}

void lpt_psi_fields(float* S, float* const G[6], int n, double boxsize, const double* d_amp) {
  fft_r2c_3d(S, n);
  lpt_psi_ij(S, G, n, boxsize, d_amp);
  for (int i = 0; i < 6; ++i) fft_c2r_3d(G[i], n);
}

void lpt_displacements(float* S, float* const G[6], int n, double boxsize, const double* d_amp) {
  lpt_psi_fields(S, G, n, boxsize, d_amp);
  const float* cg[6] = {G[0], G[1], G[2], G[3], G[4], G[5]};
  lpt_div2(G[0], cg, n);
  fft_r2c_3d(G[0], n);
  float* o5[5] = {G[1], G[2], G[3], G[4], G[5]};
  lpt_psi12(S, G[0], o5, n, boxsize, d_amp);
  fft_c2r_3d(G[1], n);
  fft_c2r_3d(G[2], n);
  fft_c2r_3d(S, n);
  fft_c2r_3d(G[3], n);
  fft_c2r_3d(G[4], n);
  fft_c2r_3d(G[5], n);
}

}  // namespace fs2

using namespace fs2;

extern "C" {

// ---- library --------------------------------------------------------------------------------------
const char* fs2_last_error(void) { return g_err.c_str(); }
int fs2_version(void) { return 1; }
int fs2_set_loglevel(int level) {
  g_loglevel = level;
  return FS2_OK;
}
int fs2_device_init(void) { return guard([] { device_init(); }); }
int fs2_synchronize(void) { return guard([] { sync(); }); }
int fs2_get_stream(void** s) {
  return guard([&] { *s = (void*)stream(); });
}

// ---- cosmology ------------------------------------------------------------------------------------
int fs2_cosmology_init(double omega_m) { return guard([&] { host::cosmology_init(omega_m); }); }
int fs2_cosmology_D_growth(double a, double* out) {
  return guard([&] { *out = host::D_growth(a); });
}
int fs2_cosmology_D2_growth(double a, double* out) {
  return guard([&] { *out = host::D2_growth(a, host::D_growth(a)); });
}
int fs2_cosmology_f_growth_rate(double a, double* out) {
  return guard([&] { *out = host::f_growth_rate(a); });
}
int fs2_cosmology_Dv_growth(double a, double* out) {
  return guard([&] { *out = host::Dv_growth(a, host::D_growth(a)); });
}
int fs2_cosmology_hubble(double a, double* out) {
  return guard([&] {
    (void)host::omega_m();
    *out = host::hubble(a);
  });
}

int fs2_cola_kick_coefficients(double a_v, double a_x, double a_vel, double* kick_factor, double* q1,
                               double* q2) {
  return guard([&] {
    // py/cola.cpp:57-66
    const double nLPT = -2.5;
    const double om = host::omega_m();
    const double ai = a_v, a = a_x, af = a_vel;
    *kick_factor = (std::pow(af, nLPT) - std::pow(ai, nLPT)) /
                   (nLPT * std::pow(a, nLPT) * std::sqrt(om / a + (1.0 - om) * a * a));
    const double g1 = host::D_growth(a);
    const double g2 = host::D2_growth(a, g1);
    *q1 = g1;
    *q2 = g2 - g1 * g1;
  });
}

int fs2_cola_drift_coefficients(double a_x, double a_v, double a_pos, double* dt, double* da1,
                                double* da2) {
  return guard([&] {
    // py/cola.cpp:102-109
    *dt = host::drift_Sq(a_x, a_pos, a_v);
    const double gi = host::D_growth(a_x), gf = host::D_growth(a_pos);
    *da1 = gf - gi;
    *da2 = host::D2_growth(a_pos, gf) - host::D2_growth(a_x, gi);
  });
}

// ---- power --------------------------------------------------------------------------------------------
int fs2_power_alloc(const char* filename, fs2_power** out) {
  return guard([&] { *out = power_load(filename); });
}
int fs2_power_alloc_eh98(double om, double ob, double h, double ns, double sigma8, fs2_power** out) {
  return guard([&] { *out = power_eh98(om, ob, h, ns, sigma8); });
}
void fs2_power_free(fs2_power* ps) { delete ps; }
int fs2_power_n(const fs2_power* ps, int* n) {
  return guard([&] { *n = ps->kind == 0 ? (int)ps->logk.size() : 0; });
}
int fs2_power_i(const fs2_power* ps, int i, double* k, double* P) {
  return guard([&] {
    if (ps->kind != 0 || i < 0 || i >= (int)ps->logk.size())
      fail(FS2_ERR_INDEX, "power spectrum index %d out of range", i);
    *k = std::exp(ps->logk[i]);
    *P = std::exp(ps->logP[i]);
  });
}
int fs2_power_P(const fs2_power* ps, double k, double* P) {
  return guard([&] { *P = power_P(ps, k); });
}
int fs2_power_sigma(const fs2_power* ps, double R, double* sigma) {
  return guard([&] { *sigma = power_sigma(ps, R); });
}

// ---- particles ----------------------------------------------------------------------------------------
int fs2_particles_alloc(uint64_t np_alloc, double boxsize, fs2_particles** out) {
  return guard([&] {
    fs2_particles* p = new fs2_particles();
    p->np_alloc = np_alloc;
    p->boxsize = boxsize;
    try {
      const size_t nf = (size_t)15 * std::max<uint64_t>(np_alloc, 1);
      p->planes = (float*)dmalloc(nf * sizeof(float));
      p->id = (uint64_t*)dmalloc(std::max<uint64_t>(np_alloc, 1) * sizeof(uint64_t));
      FS2_HIP(hipMemsetAsync(p->planes, 0, nf * sizeof(float), stream()));
      FS2_HIP(hipMemsetAsync(p->id, 0, std::max<uint64_t>(np_alloc, 1) * sizeof(uint64_t), stream()));
    } catch (...) {
      dfree(p->planes);
      dfree(p->id);
      delete p;
      throw;
    }
    *out = p;
  });
}

void fs2_particles_free(fs2_particles* p) {
  if (!p) return;
  if (g_inited) (void)hipStreamSynchronize(g_stream);
  dfree(p->planes);
  dfree(p->id);
  delete p;
}

int fs2_particles_len(const fs2_particles* p, uint64_t* n) {
  return guard([&] { *n = p->np_local; });
}
int fs2_particles_np_total(const fs2_particles* p, uint64_t* n) {
  return guard([&] { *n = p->np_total; });
}
int fs2_particles_update_np_total(fs2_particles* p) {
  return guard([&] { p->np_total = p->np_local; });
}
int fs2_particles_np_allocated(const fs2_particles* p, uint64_t* n) {
  return guard([&] { *n = p->np_alloc; });
}
int fs2_particles_boxsize(const fs2_particles* p, double* b) {
  return guard([&] { *b = p->boxsize; });
}

int fs2_particles_append(fs2_particles* p, const double* x, uint64_t n, int64_t row_stride,
                         int64_t col_stride) {
  return guard([&] {
    check_p(p);
    if (n == 0) {
      p->np_total = p->np_local;
      return;
    }
    if (p->np_local + n > p->np_alloc)
      fail(FS2_ERR_MEMORY, "append of %llu particles exceeds np_alloc %llu (have %llu)",
           (unsigned long long)n, (unsigned long long)p->np_alloc,
           (unsigned long long)p->np_local);
    flush_wrap(p);
    std::vector<float> tmp(3 * n);
    for (uint64_t i = 0; i < n; ++i)
      for (int a = 0; a < 3; ++a) tmp[a * n + i] = (float)x[i * row_stride + a * col_stride];
    for (int a = 0; a < 3; ++a)
      FS2_HIP(hipMemcpyAsync(p->x() + a * p->np_alloc + p->np_local, tmp.data() + a * n,
                             n * sizeof(float), hipMemcpyHostToDevice, stream()));
    sync();
    p->np_local += n;
    p->np_total = p->np_local;
    touch_x(p);
  });
}

int fs2_particles_resize(fs2_particles* p, uint64_t n) {
  return guard([&] {
    check_p(p);
    if (n > p->np_alloc)
      fail(FS2_ERR_MEMORY, "resize to %llu particles exceeds np_alloc %llu", (unsigned long long)n,
           (unsigned long long)p->np_alloc);
    flush_wrap(p);
    p->np_local = n;
    p->np_total = n;
    touch_x(p);
  });
}

int fs2_particles_clear(fs2_particles* p) {
  return guard([&] {
    check_p(p);
    p->wrap_pending = false;
    p->np_local = 0;
    p->np_total = 0;
    touch_x(p);
  });
}

int fs2_particles_get_field(fs2_particles* p, int field, float* out) {
  return guard([&] {
    check_p(p);
    const float* d = nullptr;
    switch (field) {
      case FS2_FIELD_X: flush_wrap(p); d = p->x(); break;
      case FS2_FIELD_DX1: d = p->dx1(); break;
      case FS2_FIELD_DX2: d = p->dx2(); break;
      case FS2_FIELD_V: d = p->v(); break;
      case FS2_FIELD_FORCE: d = p->f(); break;
      default: fail(FS2_ERR_VALUE, "unknown particle field %d", field);
    }
    download_rows(d, p->np_alloc, p->np_local, out);
  });
}

int fs2_particles_set_field(fs2_particles* p, int field, const float* in) {
  return guard([&] {
    check_p(p);
    float* d = nullptr;
    switch (field) {
      case FS2_FIELD_X: flush_wrap(p); d = p->x(); break;
      case FS2_FIELD_DX1: d = p->dx1(); break;
      case FS2_FIELD_DX2: d = p->dx2(); break;
      case FS2_FIELD_V: d = p->v(); break;
      case FS2_FIELD_FORCE: d = p->f(); break;
      default: fail(FS2_ERR_VALUE, "unknown particle field %d", field);
    }
    upload_rows(d, p->np_alloc, p->np_local, in);
    if (field == FS2_FIELD_X) touch_x(p);
  });
}

int fs2_particles_get_id(fs2_particles* p, uint64_t* out) {
  return guard([&] {
    check_p(p);
    if (p->np_local == 0) return;
    FS2_HIP(hipMemcpyAsync(out, p->id, p->np_local * sizeof(uint64_t), hipMemcpyDeviceToHost,
                           stream()));
    sync();
  });
}

int fs2_particles_get_records(fs2_particles* p, uint64_t start, uint64_t count, void* out) {
  return guard([&] {
    check_p(p);
    if (start > p->np_local || count > p->np_local - start)
      fail(FS2_ERR_INDEX, "record range [%llu, %llu) outside [0, %llu)",
           (unsigned long long)start, (unsigned long long)(start + count),
           (unsigned long long)p->np_local);
    if (count == 0) return;
    flush_wrap(p);
    // py/particle.h:12-18: x v dx1 dx2 (float[3] each) + u64 id = 56 bytes
    std::vector<float> f(12 * count);
    std::vector<uint64_t> ids(count);
    const float* src[4] = {p->x(), p->v(), p->dx1(), p->dx2()};
    for (int k = 0; k < 4; ++k)
      for (int a = 0; a < 3; ++a)
        FS2_HIP(hipMemcpyAsync(f.data() + (3 * k + a) * count, src[k] + a * p->np_alloc + start,
                               count * sizeof(float), hipMemcpyDeviceToHost, stream()));
    FS2_HIP(hipMemcpyAsync(ids.data(), p->id + start, count * sizeof(uint64_t),
                           hipMemcpyDeviceToHost, stream()));
    sync();
    unsigned char* o = (unsigned char*)out;
    for (uint64_t i = 0; i < count; ++i) {
      float rec[12];
      for (int c = 0; c < 12; ++c) rec[c] = f[c * count + i];
      std::memcpy(o + 56 * i, rec, 48);
      std::memcpy(o + 56 * i + 48, &ids[i], 8);
    }
  });
}

int fs2_particles_periodic_wrapup(fs2_particles* p) {
  return guard([&] {
    check_p(p);
    p->wrap_pending = false;
    wrap_positions(pos_of(p), (float)p->boxsize);
    touch_x(p);
  });
}

int fs2_particles_get_times(const fs2_particles* p, double* a_x, double* a_v, double* a_f) {
  return guard([&] {
    *a_x = p->a_x;
    *a_v = p->a_v;
    *a_f = p->a_f;
  });
}

int fs2_particles_set_times(fs2_particles* p, double a_x, double a_v, double a_f) {
  return guard([&] {
    p->a_x = a_x;
    p->a_v = a_v;
    p->a_f = a_f;
  });
}

int fs2_particles_device_fields(fs2_particles* p, float** x, float** v, float** dx1, float** dx2,
                                float** force, uint64_t** id, uint64_t* stride) {
  return guard([&] {
    check_p(p);
    flush_wrap(p);
    if (x) *x = p->x();
    if (v) *v = p->v();
    if (dx1) *dx1 = p->dx1();
    if (dx2) *dx2 = p->dx2();
    if (force) *force = p->f();
    if (id) *id = p->id;
    if (stride) *stride = p->np_alloc;
    touch_x(p);   // the caller may write the planes
    p->planes_exposed = true;
  });
}

int fs2_particles_sync(fs2_particles* p) {
  return guard([&] { check_p(p); });
}

int fs2_particles_copy(fs2_particles* dst, const fs2_particles* src) {
  return guard([&] {
    if (!dst || !src) fail(FS2_ERR_VALUE, "null particles handle");
    if (dst->np_alloc != src->np_alloc) fail(FS2_ERR_VALUE, "np_alloc mismatch in copy");
    FS2_HIP(hipMemcpyAsync(dst->planes, src->planes, (size_t)15 * src->np_alloc * sizeof(float),
                           hipMemcpyDeviceToDevice, stream()));
    FS2_HIP(hipMemcpyAsync(dst->id, src->id, src->np_alloc * sizeof(uint64_t),
                           hipMemcpyDeviceToDevice, stream()));
    dst->np_local = src->np_local;
    dst->np_total = src->np_total;
    dst->boxsize = src->boxsize;
    dst->a_x = src->a_x;
    dst->a_v = src->a_v;
    dst->a_f = src->a_f;
    dst->wrap_pending = src->wrap_pending;
    dst->kick = src->kick;   // the copied planes still owe src's pending kick
    dst->drift = src->drift; // and drift
    touch_x(dst);
    if (dst->planes_exposed) {   // planes handed out to the caller hold every update
      apply_pending(dst);        // (a pending drift predates a pending wrap)
      flush_wrap(dst);
    }
  });
}

namespace {
float pm_fac(const fs2_pm* pm) {   // pm.cpp:78
  return (float)(pm->pm_factor * pm->pm_factor * pm->pm_factor);
}

// Writes the density cells a compute_density left pending (same kernel and
// records, so the same values as painting at once).
void pm_materialize(fs2_pm* pm) {
  if (!pm->density_pending) return;
  pm->density_pending = false;
  paint_tiled_cells((int64_t)pm->binned_np, pm->mesh, pm->nc, pm->boxsize_f, pm_fac(pm), true,
                    pm->paint, pm->tiles);
}

void grid_sync(fs2_grid* g) {
  if (g->owner) pm_materialize(g->owner);
}
}  // namespace

// ---- grids -------------------------------------------------------------------------------------------------
int fs2_grid_alloc(int nc, fs2_grid** out) {
  return guard([&] { *out = grid_new(nc); });
}
void fs2_grid_free(fs2_grid* g) {
  if (!g || !g->owning) return;
  if (g_inited) (void)hipStreamSynchronize(g_stream);
  dfree(g->d);
  delete g;
}
int fs2_grid_nc(const fs2_grid* g, int* nc) {
  return guard([&] { *nc = g->nc; });
}
int fs2_grid_set_test_data(fs2_grid* g) {
  return guard([&] {
    grid_sync(g);
    // py/py_fft.cpp:57-85: fx[ix,iy,iz] = (ix nc + iy) nc + iz + 1
    const int n = g->nc;
    std::vector<float> h((size_t)n * n * n);
    for (size_t i = 0; i < h.size(); ++i) h[i] = (float)(i + 1);
    grid_set_real(g, h.data());
  });
}
int fs2_grid_get_real(fs2_grid* g, float* out) {
  return guard([&] { grid_sync(g); grid_get_real(g, out); });
}
int fs2_grid_set_real(fs2_grid* g, const float* in) {
  return guard([&] { grid_sync(g); grid_set_real(g, in); });
}
int fs2_grid_get_complex(fs2_grid* g, float* out) {
  return guard([&] {
    grid_sync(g);
    const int n = g->nc;
    const size_t w = (size_t)(n / 2 + 1) * 8;
    FS2_HIP(hipMemcpy2DAsync(out, w, g->d, (size_t)grid_cpitch(n) * 8, w, (size_t)n * n,
                             hipMemcpyDeviceToHost, stream()));
    sync();
  });
}
int fs2_grid_set_complex(fs2_grid* g, const float* in) {
  return guard([&] {
    grid_sync(g);
    const int n = g->nc;
    const size_t w = (size_t)(n / 2 + 1) * 8;
    FS2_HIP(hipMemcpy2DAsync(g->d, (size_t)grid_cpitch(n) * 8, in, w, w, (size_t)n * n,
                             hipMemcpyHostToDevice, stream()));
    sync();
  });
}
int fs2_grid_r2c(fs2_grid* g) {
  return guard([&] { grid_sync(g); fft_r2c_3d(g->d, g->nc); });
}
int fs2_grid_c2r(fs2_grid* g) {
  return guard([&] { grid_sync(g); fft_c2r_3d(g->d, g->nc); });
}

// ---- PM ---------------------------------------------------------------------------------------------------------
int fs2_pm_create(int nc_pm, double pm_factor, double boxsize, fs2_pm** out) {
  return guard([&] {
    // py/pm.cpp:245-255
    if (nc_pm <= 1) fail(FS2_ERR_RUNTIME, "nc_pm (= %d) must be larger than 1", nc_pm);
    if (boxsize <= 0) fail(FS2_ERR_RUNTIME, "boxsize (= %f) must be positive", boxsize);
    if (!fft_supported(nc_pm)) fail(FS2_ERR_VALUE, "nc_pm = %d is not a supported FFT size", nc_pm);
    fs2_pm* pm = new fs2_pm();
    pm->nc = nc_pm;
    pm->pm_factor = pm_factor;
    pm->boxsize = boxsize;
    pm->boxsize_f = (float)boxsize;
    const size_t bytes = grid_floats(nc_pm) * sizeof(float);
    try {
      pm->mesh = (float*)dmalloc(bytes);
      for (int a = 0; a < 3; ++a) pm->force[a] = (float*)dmalloc(bytes);
      fft_prepare(nc_pm);   // the mesh's transforms never allocate after this
    } catch (...) {
      dfree(pm->mesh);
      for (int a = 0; a < 3; ++a) dfree(pm->force[a]);
      delete pm;
      throw;
    }
    pm->tiles = column_tile_opts(nc_pm);
    pm->density.nc = nc_pm;
    pm->density.d = pm->mesh;
    pm->density.owning = false;
    pm->density.owner = pm;
    *out = pm;
  });
}

void fs2_pm_destroy(fs2_pm* pm) {
  if (!pm) return;
  if (g_inited) (void)hipStreamSynchronize(g_stream);
  free_paint_scratch(pm->paint);
  dfree(pm->mesh);
  for (int a = 0; a < 3; ++a) dfree(pm->force[a]);
  delete pm;
}

int fs2_pm_params(const fs2_pm* pm, int* nc, double* pm_factor, double* boxsize) {
  return guard([&] {
    *nc = pm->nc;
    *pm_factor = pm->pm_factor;
    *boxsize = pm->boxsize;
  });
}

int fs2_pm_set_paint_variant(fs2_pm* pm, int variant) {
  return guard([&] {
    if (variant != FS2_PAINT_ATOMIC && variant != FS2_PAINT_SORTED && variant != FS2_PAINT_GLOBAL)
      fail(FS2_ERR_VALUE, "paint variant %d not available", variant);
    pm->variant = variant;
  });
}

int fs2_pm_send_positions(fs2_pm* pm, fs2_particles* p) {
  return guard([&] {
    (void)pm;
    // only marks the wrap: a pending kick and drift stay pending for the
    // density binning, which applies them before it wraps.  Planes handed
    // out to the caller are brought up to date at once.
    if (!p) fail(FS2_ERR_VALUE, "null particles handle");
    p->wrap_pending = true;
    if (p->planes_exposed) {
      apply_pending(p);
      flush_wrap(p);
    }
  });
}

int fs2_pm_compute_density(fs2_pm* pm, fs2_particles* p, fs2_grid** density) {
  return guard([&] {
    if (!p) fail(FS2_ERR_VALUE, "null particles handle");
    // a pending COLA step runs inside the tiled paint's binning pass
    const bool step_in_bin = p->drift.pending && pm->variant != FS2_PAINT_GLOBAL;
    if (!step_in_bin) check_p(p);
    const int n = pm->nc;
    const float fac = pm_fac(pm);
    pm->density_pending = false;
    if (pm->variant == FS2_PAINT_GLOBAL) {
      // pm.cpp:414-426 clear to -1, then the scatter of pm.cpp:57-139
      fill_grid(pm->mesh, n, -1.0f);
      paint_atomic(pos_of(p), pm->mesh, n, pm->boxsize_f, fac, p->wrap_pending);
    } else {
      const bool det = pm->variant == FS2_PAINT_SORTED;
      ensure_paint_scratch(pm->paint, p->np_local, paint_tiled_bins(n, pm->tiles));
      if (step_in_bin) {
        paint_tiled_bin_kick_drift(pos_of(p), p->v(), p->f(), p->dx1(), p->dx2(),
                                   pending_kick_drift(p), n, pm->boxsize_f, p->wrap_pending,
                                   pm->paint, pm->tiles);
        clear_pending(p);
      } else {
        paint_tiled_bin(pos_of(p), n, pm->boxsize_f, p->wrap_pending, pm->paint, pm->tiles);
      }
      // the cells wait for compute_force to fuse them with the z r2c
      pm->density_pending = det && paint_r2c_supported(n, pm->tiles);
      if (!pm->density_pending)
        paint_tiled_cells((int64_t)p->np_local, pm->mesh, n, pm->boxsize_f, fac, det, pm->paint,
                          pm->tiles);
      if (p->wrap_pending) touch_x(p);   // the binning kernel wrapped x in place
    }
    pm->binned_version = pm->variant == FS2_PAINT_GLOBAL ? 0 : p->x_version;
    pm->binned_np = p->np_local;
    p->wrap_pending = false;
    pm->mesh_is_k = false;
    pm->status = PM_DENSITY_DONE;
    if (density) *density = &pm->density;
  });
}

int fs2_pm_check_total_density(fs2_pm* pm, double* sum) {
  return guard([&] {
    if (pm->mesh_is_k || pm->status != PM_DENSITY_DONE)
      fail(FS2_ERR_RUNTIME, "density mesh not available for the total-density check");
    pm_materialize(pm);
    const double s = grid_sum(pm->mesh, pm->nc);
    if (sum) *sum = s;
    const double n = pm->nc;
    const double tol = 10.0 * 1.1920928955078125e-07 * n * n * n;   // pm.cpp:376
    if (std::fabs(s) > tol)
      fail(FS2_ERR_ASSERTION, "total CIC density error is too large: %le > %le", s, tol);
  });
}

int fs2_pm_compute_force(fs2_pm* pm, fs2_particles* p) {
  return guard([&] {
    check_p(p);
    if (p->a_f == p->a_x) return;   // pm.cpp:311-312
    if (pm->status != PM_DENSITY_DONE) fail(FS2_ERR_RUNTIME, "PM density not ready");
    const int n = pm->nc;
    // pm.cpp:472: f1 = -1/n^3 / (2 pi / boxsize), boxsize held as a Float
    const float f1 = (float)(-1.0 / std::pow((double)n, 3.0) / (2.0 * M_PI / (double)pm->boxsize_f));
    // The paint's bin records hold the positions it binned.  They are reused
    // only when x cannot have changed since: not after the device planes were
    // handed out (fs2_particles_device_fields), whose writes bump no version.
    flush_wrap(p);
    const bool reuse_bins = pm->binned_version != 0 && pm->binned_version == p->x_version &&
                            pm->binned_np == p->np_local && !p->planes_exposed;
    pm->mesh_is_k = false;
    pm->status = PM_FORCE_DONE;   // the density mesh no longer holds delta
    if (reuse_bins && column_readout_supported(n)) {
      // the density's tile kernel fused with its z r2c, and the z c2r of the
      // three force meshes fused into the readout
      const bool z_done = pm->density_pending;
      if (z_done) {
        paint_tiled_r2c((int64_t)pm->binned_np, pm->mesh, n, pm->boxsize_f, pm_fac(pm),
                        pm->paint, pm->tiles);
        pm->density_pending = false;
      }
      fft_force_spectra(pm->mesh, pm->force, n, f1, z_done);
      readout_columns(pm->force[0], pm->force[1], pm->force[2], p->f(), (int64_t)p->np_alloc,
                      (int64_t)p->np_local, n, pm->boxsize_f, pm->paint, pm->tiles);
      pm->force_spectral = true;
    } else {
      pm_materialize(pm);
      fft_force_meshes(pm->mesh, pm->force, n, f1);
      pm->force_spectral = false;
      if (reuse_bins) {
        readout_tiled(pm->force[0], pm->force[1], pm->force[2], p->f(), (int64_t)p->np_alloc,
                      (int64_t)p->np_local, n, pm->boxsize_f, pm->paint, pm->tiles);
      } else {
        readout3(pos_of(p), pm->force[0], pm->force[1], pm->force[2], p->f(), n, pm->boxsize_f);
      }
    }
    p->a_f = p->a_x;
  });
}

int fs2_pm_get_forces(fs2_pm* pm, fs2_particles* p) {
  return guard([&] {
    (void)pm;
    check_p(p);
  });
}

int fs2_pm_get_force_mesh(fs2_pm* pm, int axis, float* out) {
  return guard([&] {
    if (axis < 0 || axis > 2) fail(FS2_ERR_VALUE, "axis %d", axis);
    if (pm->force_spectral) {   // finish the spectra the fused readout left
      fft_force_finish(pm->force, pm->nc);
      pm->force_spectral = false;
    }
    fs2_grid view;
    view.nc = pm->nc;
    view.d = pm->force[axis];
    view.owning = false;
    grid_get_real(&view, out);
  });
}

int fs2_pm_power_spectrum(fs2_pm* pm, int nbins, double dk_bin, double* k, double* P,
                          double* nmodes) {
  return guard([&] {
    if (pm->status != PM_DENSITY_DONE && !pm->mesh_is_k)
      fail(FS2_ERR_RUNTIME, "no density mesh for the power spectrum");
    if (!pm->mesh_is_k) {
      pm_materialize(pm);
      fft_r2c_3d(pm->mesh, pm->nc);
      pm->mesh_is_k = true;
    }
    power_spectrum(pm->mesh, pm->nc, pm->boxsize, nbins, dk_bin, k, P, nmodes);
  });
}

int fs2_pm_power_spectrum_noalloc(fs2_pm* pm, int nbins, double dk_bin, double* k, double* P,
                                  double* nmodes) {
  return guard([&] {
    if (nbins > FS2_PK_NOALLOC_MAX_BINS)
      fail(FS2_ERR_VALUE, "%d P(k) bins exceed the fixed scratch of %d bins", nbins,
           FS2_PK_NOALLOC_MAX_BINS);
    if (pm->status != PM_DENSITY_DONE && !pm->mesh_is_k)
      fail(FS2_ERR_RUNTIME, "no density mesh for the power spectrum");
    if (!pm->mesh_is_k) {
      pm_materialize(pm);   // kernels only: no allocation
      fft_r2c_3d(pm->mesh, pm->nc);   // twiddles built by fs2_pm_create
      pm->mesh_is_k = true;
    }
    power_spectrum_noalloc(pm->mesh, pm->nc, pm->boxsize, nbins, dk_bin, k, P, nmodes);
  });
}

// ---- COLA ---------------------------------------------------------------------------------------------------------
// The kick and the drift are recorded.  The next density binning runs both
// in its own pass over the particles (paint_tiled_bin_kick_drift); any
// other access first runs them as one cola_kick_drift pass.  Either way
// the same two updates run in order, so the result is bitwise that of the
// kick and then the drift.
int fs2_cola_kick(fs2_particles* p, double a_vel) {
  return guard([&] {
    check_p(p);   // applies an earlier pending kick
    double kf, q1, q2;
    int rc = fs2_cola_kick_coefficients(p->a_v, p->a_x, a_vel, &kf, &q1, &q2);
    if (rc) throw Error{rc, g_err};
    // cola.cpp:57, 65-66: kick_factor, q1, q2 stored as Float
    p->kick.m15om = -1.5 * host::omega_m();
    p->kick.kf = (float)kf;
    p->kick.q1 = (float)q1;
    p->kick.q2 = (float)q2;
    p->kick.pending = true;
    if (p->planes_exposed) apply_kick(p);
    p->a_v = a_vel;
  });
}

int fs2_cola_drift(fs2_particles* p, double a_pos) {
  return guard([&] {
    if (!p) fail(FS2_ERR_VALUE, "null particles handle");
    double dt, da1, da2;
    int rc = fs2_cola_drift_coefficients(p->a_x, p->a_v, a_pos, &dt, &da1, &da2);
    if (rc) throw Error{rc, g_err};
    if (p->drift.pending) apply_pending(p);   // an earlier drift (and its kick) first
    flush_wrap(p);
    // cola.cpp:106-109: da1, da2 stored as Float; dt stays double.  The
    // update runs with the next density binning (or the next other access),
    // fused with the pending kick when there is one.
    p->drift.dt = dt;
    p->drift.da1 = (float)da1;
    p->drift.da2 = (float)da2;
    p->drift.pending = true;
    if (p->planes_exposed) apply_pending(p);
    touch_x(p);
    p->a_x = a_pos;
  });
}

// ---- friends-of-friends halo finder (py/fof.cpp, py/kdtree.cpp) ---------------------------
namespace {
// Result of the last fs2_fof_find_groups (the reference keeps the group
// array and the tree in file-scope state as well, fof.cpp:10-17).
FofResult g_fof;
bool g_fof_done = false;

struct PinnedHost {
  void* p = nullptr;
  explicit PinnedHost(size_t bytes) {
    device_init();
    FS2_HIP(hipHostMalloc(&p, bytes ? bytes : 8, hipHostMallocDefault));
  }
  ~PinnedHost() {
    if (p) (void)hipHostFree(p);
  }
};
struct DeviceTmp {
  void* p;
  explicit DeviceTmp(size_t bytes) : p(dmalloc(bytes)) {}
  ~DeviceTmp() { dfree(p); }
};
}  // namespace

int fs2_fof_find_groups(fs2_particles* p, double linking_length, const double* boxsize3,
                        int quota, uint64_t* ngroups) {
  return guard([&] {
    check_p(p);
    flush_wrap(p);
    if (!(linking_length >= 0.0)) fail(FS2_ERR_VALUE, "linking length must be >= 0");
    const int64_t n = (int64_t)p->np_local;
    // final positions to the host through pinned memory
    PinnedHost h((size_t)n * 3 * sizeof(float) + (size_t)n * sizeof(int32_t));
    float* hx = (float*)h.p;
    for (int a = 0; a < 3; ++a)
      if (n > 0)
        FS2_HIP(hipMemcpyAsync(hx + a * n, p->x() + a * p->np_alloc, (size_t)n * sizeof(float),
                               hipMemcpyDeviceToHost, stream()));
    sync();
    g_fof_done = false;
    const float* xyz[3] = {hx, hx + n, hx + 2 * n};
    fof_groups(xyz, n, (int64_t)p->np_alloc, p->boxsize, linking_length, boxsize3, quota, g_fof);
    g_fof_done = true;
    // the tree build reorders x v dx1 dx2 id (kdtree.cpp:94-105 sorts the
    // particle records); the separate force array keeps its order
    if (n > 0) {
      int32_t* hperm = (int32_t*)(hx + 3 * n);
      std::copy(g_fof.perm.begin(), g_fof.perm.end(), hperm);
      DeviceTmp perm((size_t)n * sizeof(int32_t)), tmp((size_t)n * sizeof(uint64_t));
      FS2_HIP(hipMemcpyAsync(perm.p, hperm, (size_t)n * sizeof(int32_t), hipMemcpyHostToDevice,
                             stream()));
      for (int k = 0; k < 4; ++k)
        for (int a = 0; a < 3; ++a) {
          uint32_t* plane = (uint32_t*)(p->field(k) + a * p->np_alloc);
          gather_u32((uint32_t*)tmp.p, plane, (const int32_t*)perm.p, n);
          FS2_HIP(hipMemcpyAsync(plane, tmp.p, (size_t)n * sizeof(float),
                                 hipMemcpyDeviceToDevice, stream()));
        }
      gather_u64((uint64_t*)tmp.p, p->id, (const int32_t*)perm.p, n);
      FS2_HIP(hipMemcpyAsync(p->id, tmp.p, (size_t)n * sizeof(uint64_t), hipMemcpyDeviceToDevice,
                             stream()));
      sync();   // the temporaries and the pinned buffer are freed on return
    }
    touch_x(p);
    if (ngroups) {
      uint64_t c = 0;
      for (int64_t i = 0; i < n; ++i) c += g_fof.grp[i] == (int32_t)i;
      *ngroups = c;
    }
  });
}

int fs2_fof_grp(int32_t* out, uint64_t* n) {
  return guard([&] {
    if (n) *n = g_fof_done ? g_fof.grp.size() : 0;
    if (out && g_fof_done) std::copy(g_fof.grp.begin(), g_fof.grp.end(), out);
  });
}

int fs2_fof_nfof(int32_t* out, uint64_t* n) {
  return guard([&] {
    if (n) *n = g_fof_done ? g_fof.grp.size() : 0;
    if (out && g_fof_done) {
      std::vector<int32_t> c;
      fof_counts(g_fof.grp, c);
      std::copy(c.begin(), c.end(), out);
    }
  });
}

int fs2_kdtree_height(uint64_t* height) {
  return guard([&] {
    if (!g_fof_done) fail(FS2_ERR_RUNTIME, "no kd-tree: run fs2_fof_find_groups first");
    *height = (uint64_t)g_fof.height;
  });
}

int fs2_kdtree_nodes(fs2_kdtree_node* out, uint64_t* n) {
  return guard([&] {
    if (!g_fof_done) fail(FS2_ERR_RUNTIME, "no kd-tree: run fs2_fof_find_groups first");
    if (n) *n = g_fof.tree.size();
    if (out)
      for (size_t i = 0; i < g_fof.tree.size(); ++i) {
        const KdNode& t = g_fof.tree[i];
        out[i] = fs2_kdtree_node{t.k, t.left, t.right, t.ibegin, t.iend};
      }
  });
}

// ---- output: Gadget-1 binary ---------------------------------------------------------------
namespace {
// py/gadget_file.h:6-25: the 256-byte Gadget-1 header (layout kept for the format)
struct GadgetHeader {
  int32_t np[6];
  double mass[6];
  double time;
  double redshift;
  int32_t flag_sfr;
  int32_t flag_feedback;
  uint32_t np_total[6];
  int32_t flag_cooling;
  int32_t num_files;
  double boxsize;
  double omega0;
  double omega_lambda;
  double hubble_param;
  int32_t flag_stellarage;
  int32_t flag_metals;
  uint32_t np_total_highword[6];
  int32_t flag_entropy_instead_u;
  char fill[60];
};
static_assert(sizeof(GadgetHeader) == 256, "Gadget-1 header is 256 bytes");

// py/const.h:12-23: internal units 1/h Mpc, km/s, M_sun; rho_crit_0 = 3 H0^2/(8 pi G)
double rho_crit_0() {
  const double unit_length = 3.085678e24, unit_mass = 1.989e33, unit_velocity = 1.0e5;
  const double G_cgs = 6.672e-8;
  const double unit_time = unit_length / unit_velocity;
  const double G = G_cgs * unit_mass * (unit_time * unit_time) /
                   (unit_length * unit_length * unit_length);
  const double H0 = 100.0;
  return 3.0 * H0 * H0 / (8.0 * M_PI * G);
}

void put_block(std::FILE* fp, const void* data, size_t bytes, const char* path) {
  const int32_t len = (int32_t)bytes;
  if (std::fwrite(&len, 4, 1, fp) != 1 || (bytes && std::fwrite(data, 1, bytes, fp) != bytes) ||
      std::fwrite(&len, 4, 1, fp) != 1)
    fail(FS2_ERR_IO, "write to %s failed", path);
}
}  // namespace

int fs2_cosmology_omega_m(double* out) {
  return guard([&] { *out = host::omega_m(); });
}

int fs2_write_gadget(fs2_particles* p, const char* filename, int use_long_id, double h) {
  return guard([&] {
    check_p(p);
    if (!filename) fail(FS2_ERR_VALUE, "null file name");
    const uint64_t np = p->np_local;
    // positions, velocities / sqrt(a) (Gadget convention), ids (gadget_file.cpp:70-114)
    std::vector<float> x(3 * np), v(3 * np);
    std::vector<uint64_t> id(np);
    int rc = fs2_particles_get_field(p, FS2_FIELD_X, x.data());
    if (!rc) rc = fs2_particles_get_field(p, FS2_FIELD_V, v.data());
    if (!rc) rc = fs2_particles_get_id(p, id.data());
    if (rc) fail(rc, "%s", fs2_last_error());
    GadgetHeader hd;
    std::memset(&hd, 0, sizeof(hd));
    const double om = host::omega_m();
    const uint64_t ntot = np;   // comm_sum of np_local over the one rank
    hd.np[1] = (int32_t)np;
    hd.mass[1] = om * rho_crit_0() * std::pow(p->boxsize, 3.0) / (double)ntot;
    hd.time = p->a_x;
    hd.redshift = 1.0 / hd.time - 1.0;
    hd.np_total[1] = (uint32_t)ntot;
    hd.np_total_highword[1] = (uint32_t)(ntot >> 32);
    hd.num_files = 1;
    hd.boxsize = p->boxsize;
    hd.omega0 = om;
    hd.omega_lambda = 1.0 - om;
    hd.hubble_param = h;
    const float vfac = (float)(1.0 / std::sqrt(p->a_x));
    for (float& c : v) c = (float)(vfac * c);
    std::FILE* fp = std::fopen(filename, "wb");
    if (!fp) fail(FS2_ERR_IO, "unable to write to file: %s", filename);
    try {
      put_block(fp, &hd, sizeof(hd), filename);
      put_block(fp, x.data(), x.size() * sizeof(float), filename);
      put_block(fp, v.data(), v.size() * sizeof(float), filename);
      if (use_long_id) {
        put_block(fp, id.data(), id.size() * sizeof(uint64_t), filename);
      } else {
        std::vector<uint32_t> id32(id.begin(), id.end());
        put_block(fp, id32.data(), id32.size() * sizeof(uint32_t), filename);
      }
    } catch (...) {
      std::fclose(fp);
      throw;
    }
    if (std::fclose(fp) != 0) fail(FS2_ERR_IO, "write to %s failed", filename);
  });
}

// ---- standard leapfrog --------------------------------------------------------------------------
int fs2_leapfrog_factors(double ai, double af, double* kick_factor, double* dt) {
  return guard([&] {
    if (kick_factor) *kick_factor = host::leapfrog_Sphi(ai, af);   // leapfrog.cpp:96-115
    if (dt) *dt = host::leapfrog_Sq(ai, af);                          // leapfrog.cpp:117-141
  });
}

int fs2_leapfrog_initial_velocity(fs2_particles* p, double a) {
  return guard([&] {
    check_p(p);
    // leapfrog.cpp:22-28: Float da1, da2, Dv, D2v (D2v from da2 here)
    const float da1 = (float)host::D_growth(a);
    const float da2 = (float)host::D2_growth(a, (double)da1);
    const float Dv = (float)host::Dv_growth(a, (double)da1);
    const float D2v = (float)host::D2v_growth(a, (double)da2);
    leapfrog_initial_velocity(p->v(), p->dx1(), p->dx2(), (int64_t)p->np_alloc,
                              (int64_t)p->np_local, Dv, D2v);
    p->a_v = a;
  });
}

int fs2_leapfrog_kick(fs2_particles* p, double a_vel) {
  return guard([&] {
    check_p(p);
    const double kf = host::leapfrog_Sphi(p->a_v, a_vel);
    leapfrog_kick(p->v(), p->f(), (int64_t)p->np_alloc, (int64_t)p->np_local,
                  -1.5 * host::omega_m(), (float)kf);   // kick_factor is a Float
    p->a_v = a_vel;
  });
}

int fs2_leapfrog_drift(fs2_particles* p, double a_pos) {
  return guard([&] {
    check_p(p);
    const double dt = host::leapfrog_Sq(p->a_x, a_pos);
    flush_wrap(p);
    leapfrog_drift(p->x(), p->v(), (int64_t)p->np_alloc, (int64_t)p->np_local, dt);
    touch_x(p);
    p->a_x = a_pos;
  });
}

int fs2_cola_velocity(fs2_particles* p, float* out) {
  return guard([&] {
    check_p(p);
    const double a = p->a_x;
    const double D1 = host::D_growth(a);
    const float Dv = (float)host::Dv_growth(a, D1);
    const float D2v = (float)host::D2v_growth(a, D1);   // cola.cpp:138 passes D1
    // the kernel writes planes at stride np_alloc
    float* full = (float*)dmalloc(3 * std::max<uint64_t>(p->np_alloc, 1) * sizeof(float));
    try {
      cola_velocity(full, p->v(), p->dx1(), p->dx2(), (int64_t)p->np_alloc, (int64_t)p->np_local,
                    Dv, D2v);
      download_rows(full, p->np_alloc, p->np_local, out);
    } catch (...) {
      dfree(full);
      throw;
    }
    dfree(full);
  });
}

// ---- 2LPT ---------------------------------------------------------------------------------------------------------
int fs2_lpt_set_offset(double offset) {
  return guard([&] { lpt_offset = (float)offset; });
}

int fs2_lpt_set_zeldovich_force(fs2_particles* p, double a) {
  return guard([&] {
    check_p(p);
    const float g1 = (float)host::D_growth(a);   // py_lpt.cpp:74
    scale_field(p->f(), p->dx1(), (int64_t)p->np_alloc, (int64_t)p->np_local, g1);
  });
}

int fs2_lpt_white_noise(int nc, uint64_t seed, float* out) {
  return guard([&] {
    fs2_grid* g = grid_new(nc);
    try {
      white_noise(g->d, nc, seed);
      grid_get_real(g, out);
    } catch (...) {
      fs2_grid_free(g);
      throw;
    }
    fs2_grid_free(g);
  });
}

int fs2_lpt(int nc, double boxsize, double a, uint64_t seed, const fs2_power* ps, const char* kind,
            fs2_particles** out) {
  return guard([&] {
    if (!ps) fail(FS2_ERR_VALUE, "null power spectrum");
    const std::string k(kind ? kind : "");
    if (k != "cola" && k != "zeldovich" && k != "2lpt")
      fail(FS2_ERR_VALUE, "unknown LPT kind '%s' (zeldovich, 2lpt, cola)", k.c_str());
    if (nc < 4 || nc % 2 != 0 || !fft_supported(nc))
      fail(FS2_ERR_VALUE, "nc = %d not supported for 2LPT", nc);
    if (boxsize <= 0) fail(FS2_ERR_VALUE, "boxsize must be positive");
    (void)host::omega_m();   // cosmology must be initialised

    const std::vector<double> amp = lpt_amp_table(nc, boxsize, ps);
    const int nk2 = (int)amp.size();

    const size_t gbytes = grid_floats(nc) * sizeof(float);
    std::vector<float*> bufs;
    double* d_amp = nullptr;
    fs2_particles* p = nullptr;
    auto cleanup = [&] {
      for (float* b : bufs) dfree(b);
      dfree(d_amp);
    };
    try {
      for (int i = 0; i < 7; ++i) bufs.push_back((float*)dmalloc(gbytes));
      d_amp = (double*)dmalloc(nk2 * sizeof(double));
      FS2_HIP(hipMemcpyAsync(d_amp, amp.data(), nk2 * sizeof(double), hipMemcpyHostToDevice,
                             stream()));
      float* S = bufs[0];
      float* G[6] = {bufs[1], bufs[2], bufs[3], bufs[4], bufs[5], bufs[6]};
      white_noise(S, nc, seed);
      lpt_displacements(S, G, nc, boxsize, d_amp);

      // lpt.cpp:156-175: Float D1; D2, Dv, D2v evaluated from the Float D1
      const float D1 = (float)host::D_growth(a);
      const float D2 = k == "zeldovich" ? 0.f : (float)host::D2_growth(a, (double)D1);
      float Dv = 0.f, D2v = 0.f;
      if (k == "zeldovich") {
        Dv = (float)host::Dv_growth(a, (double)D1);
      } else if (k == "2lpt") {
        Dv = (float)host::Dv_growth(a, (double)D1);
        D2v = (float)host::D2v_growth(a, (double)D1);   // lpt.cpp:172 passes D1
      }
      const uint64_t np = (uint64_t)nc * nc * nc;
      int rc = fs2_particles_alloc(np, boxsize, &p);
      if (rc) throw Error{rc, g_err};
      LatticeArgs la{p->x(), p->v(), p->dx1(), p->dx2(), p->id, (int64_t)p->np_alloc};
      const float* psi1[3] = {G[1], G[2], S};
      const float* psi2[3] = {G[3], G[4], G[5]};
      const float dx = (float)(boxsize / nc);   // lpt.cpp:144
      lpt_lattice(la, psi1, psi2, nc, dx, lpt_offset, D1, D2, Dv, D2v,
                  1.0 / std::pow((double)nc, 3.0));
      sync();
      p->np_local = np;
      p->np_total = np;
      p->a_x = a;
      p->a_v = a;
      touch_x(p);
    } catch (...) {
      cleanup();
      fs2_particles_free(p);
      throw;
    }
    cleanup();
    *out = p;
  });
}

// ---- named device allocations (py/mem.cpp, py/py_mem.cpp) -------------------------------------
struct fs2_mem {
  std::string name;
  void* buf = nullptr;
  uint64_t size_alloc = 0, size_using = 0;
};

int fs2_mem_alloc(const char* name, uint64_t size, fs2_mem** out) {
  return guard([&] {
    if (!out) fail(FS2_ERR_VALUE, "null output");
    auto m = std::make_unique<fs2_mem>();
    m->name = name ? name : "";
    if (size > 0) {
      const uint64_t aligned = (size + 15) / 16 * 16;   // mem.cpp:48-50, ALGN 16
      m->buf = dmalloc((size_t)aligned);
      m->size_alloc = aligned;
    }
    *out = m.release();
  });
}

int fs2_mem_info(const fs2_mem* m, uint64_t* size_alloc, uint64_t* size_using) {
  return guard([&] {
    if (!m) fail(FS2_ERR_VALUE, "null mem handle");
    if (size_alloc) *size_alloc = m->size_alloc;
    if (size_using) *size_using = m->size_using;
  });
}

void fs2_mem_free(fs2_mem* m) {
  if (!m) return;
  if (m->buf) {
    if (g_inited) (void)hipStreamSynchronize(g_stream);
    dfree(m->buf);
  }
  delete m;
}

// ---- streaming-copy probe (roofline harness) ----------------------------------------------
int fs2_copy_bandwidth(uint64_t bytes, int grid, int reps, double* ms) {
  return guard([&] {
    if (!ms) fail(FS2_ERR_VALUE, "null output");
    *ms = copy_probe_ms(bytes, grid, reps);
  });
}

// ---- kernel timing ---------------------------------------------------------------------------------------------
int fs2_profile_enable(int on) {
  return guard([&] {
    if (!on) prof_resolve();
    g_prof = on != 0;
  });
}
int fs2_profile_reset(void) {
  return guard([&] {
    prof_resolve();
    g_recs.clear();
  });
}
int fs2_profile_count(int* n) {
  return guard([&] {
    prof_resolve();
    *n = (int)g_recs.size();
  });
}
int fs2_profile_get(int i, const char** name, uint64_t* launches, double* total_ms,
                    double* bytes_per_launch) {
  return guard([&] {
    prof_resolve();
    if (i < 0 || i >= (int)g_recs.size()) fail(FS2_ERR_INDEX, "profile index %d", i);
    *name = g_recs[i].name.c_str();
    *launches = g_recs[i].launches;
    *total_ms = g_recs[i].total_ms;
    *bytes_per_launch = g_recs[i].bytes;
  });
}

}  // extern "C"
