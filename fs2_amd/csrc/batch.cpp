// Batched ensemble of small COLA boxes in one hipGraph.
//
// No reference counterpart: the reference runs one box per process
// (script/cola.py:28-48).  Here many independent realisations of a few
// sizes run together.  Boxes of one size form a class whose particles sit
// box-major in one set of planes [3][nbox * np_box] and whose meshes are
// nbox grids back to back, so every kernel of the step (py/pm.cpp paint,
// FFT, Green x gradient, readout; py/cola.cpp kick, drift) covers the whole
// class in one launch.  Each class runs on its own stream; the whole
// nstep-step run of every class is captured once into a hipGraph (forked
// from and joined to the library stream) and replayed with one launch.
//
// The paint tile shape of each class is chosen by fs2_batch_autotune: every
// candidate shape is timed on the class's own data (bin + scan + scatter +
// tile paint + tiled readout, HIP events) and the fastest is kept.  The
// deterministic paint sums are integers, so the choice never changes a bit
// of the results.
#include <array>
#include <cmath>
#include <vector>

#include "fs2_internal.h"

namespace {

struct BatchClass {
  int nl = 0, nm = 0, nbox = 0;
  double boxsize = 0.0;
  float boxsize_f = 0.f, fac = 1.f, f1 = 0.f;
  int64_t npb = 0, np = 0;
  float* planes = nullptr;   // [5][3][np]: x, v, dx1, dx2, force
  float* ic = nullptr;       // [4][3][np]: x, v, dx1, dx2 after fs2_batch_init_lpt
  float* mesh = nullptr;     // nbox grids
  float* F[3] = {};          // nbox grids each
  fs2::PaintScratch paint;
  fs2::TileOpts tile;
  hipStream_t s = nullptr;
  float* field(int k) const { return planes + (size_t)k * 3 * np; }
};

}  // namespace

struct fs2_batch {
  std::vector<BatchClass> cls;
  bool have_ic = false;
  double a_ic = 0.0;
  // captured schedule
  int nstep = 0;
  double a0 = 0.0, a1 = 0.0;
  bool have_schedule = false;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  hipEvent_t fork = nullptr;
  std::vector<hipEvent_t> join;
};

namespace fs2 {
namespace {

// Tile shapes the autotuner tries (power-of-two extents; shapes that do not
// divide the mesh or exceed the LDS budget are skipped).
const int kShapes[][3] = {{8, 8, 64},  {8, 16, 32}, {8, 8, 32}, {4, 8, 64},  {16, 16, 16},
                          {8, 16, 16}, {4, 16, 32}, {8, 8, 16}, {16, 8, 32}, {4, 8, 32},
                          {16, 16, 32}};

bool shape_ok(int n, const int* t) {
  for (int a = 0; a < 3; ++a)
    if (t[a] > n || n % t[a]) return false;
  const size_t rows = (size_t)(t[0] + 1) * (t[1] + 1);
  const size_t readout_lds = 3 * rows * (t[2] + 1) * sizeof(float) + rows * sizeof(int);
  const size_t paint_lds = (size_t)t[0] * t[1] * t[2] * sizeof(unsigned long long);
  return readout_lds <= 160 * 1024 && paint_lds <= 160 * 1024;
}

size_t max_bins(const BatchClass& c) {
  size_t m = 0;
  for (const auto& t : kShapes) {
    if (!shape_ok(c.nm, t)) continue;
    TileOpts o;
    o.tx = t[0];
    o.ty = t[1];
    o.tz = t[2];
    o.nbox = c.nbox;
    m = std::max(m, paint_tiled_bins(c.nm, o));
  }
  return std::max(m, paint_tiled_bins(c.nm, c.tile));
}

void destroy(fs2_batch* b) {
  if (!b) return;
  (void)hipStreamSynchronize(library_stream());
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  if (b->fork) (void)hipEventDestroy(b->fork);
  for (hipEvent_t e : b->join) (void)hipEventDestroy(e);
  for (BatchClass& c : b->cls) {
    if (c.s) (void)hipStreamSynchronize(c.s);
    dfree(c.planes);
    dfree(c.ic);
    dfree(c.mesh);
    for (float* g : c.F) dfree(g);
    free_paint_scratch(c.paint);
    if (c.s) (void)hipStreamDestroy(c.s);
  }
  delete b;
}

void drop_graph(fs2_batch* b) {
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  b->exec = nullptr;
  b->graph = nullptr;
}

// density -> force meshes -> forces of every box of the class; with `step`,
// the previous step's kick and drift run inside the density's binning pass
// (paint_tiled_bin_kick_drift, bitwise the same as running them first)
void class_force(BatchClass& c, const KickDrift* step = nullptr) {
  const PosArgs pos{c.field(0), c.np, c.np};
  if (step) {
    paint_tiled_bin_kick_drift(pos, c.field(1), c.field(4), c.field(2), c.field(3), *step, c.nm,
                               c.boxsize_f, true, c.paint, c.tile);
    paint_tiled_cells(c.np, c.mesh, c.nm, c.boxsize_f, c.fac, true, c.paint, c.tile);
  } else {
    paint_tiled(pos, c.mesh, c.nm, c.boxsize_f, c.fac, true, true, c.paint, c.tile);
  }
  fft_force_meshes(c.mesh, c.F, c.nm, c.f1, c.nbox);
  readout_tiled(c.F[0], c.F[1], c.F[2], c.field(4), c.np, c.np, c.nm, c.boxsize_f, c.paint,
                c.tile);
}

struct StepCoef {
  float kf, q1, q2, da1, da2;
  double dt;
};

std::vector<StepCoef> schedule(double a0, double a1, int nstep) {
  std::vector<StepCoef> out;
  double a_x = a0, a_v = a0;
  for (int i = 0; i < nstep; ++i) {
    // script/cola.py:37-44
    const double a_vel = a0 + (a1 - a0) / nstep * (i + 0.5);
    const double a_pos = a0 + (a1 - a0) / nstep * (i + 1.0);
    double kf, q1, q2, dt, da1, da2;
    int rc = fs2_cola_kick_coefficients(a_v, a_x, a_vel, &kf, &q1, &q2);
    if (rc) fail(rc, "%s", fs2_last_error());
    a_v = a_vel;
    rc = fs2_cola_drift_coefficients(a_x, a_v, a_pos, &dt, &da1, &da2);
    if (rc) fail(rc, "%s", fs2_last_error());
    a_x = a_pos;
    out.push_back({(float)kf, (float)q1, (float)q2, (float)da1, (float)da2, dt});
  }
  return out;
}

// The whole run: every class on its own stream, forked from and joined to
// the library stream (recorded into the graph when capturing).
void run_schedule(fs2_batch* b) {
  const std::vector<StepCoef> coef = schedule(b->a0, b->a1, b->nstep);
  const double m15om = -1.5 * host::omega_m();
  hipStream_t main = library_stream();
  FS2_HIP(hipEventRecord(b->fork, main));
  for (size_t k = 0; k < b->cls.size(); ++k) {
    BatchClass& c = b->cls[k];
    FS2_HIP(hipStreamWaitEvent(c.s, b->fork, 0));
    StreamScope scope(c.s);
    // each step's kick and drift run in the next step's binning pass; the
    // last step's after the loop
    for (size_t i = 0; i < coef.size(); ++i) {
      if (i == 0) {
        class_force(c);
      } else {
        const StepCoef& s = coef[i - 1];
        const KickDrift kd{1, m15om, s.kf, s.q1, s.q2, s.dt, s.da1, s.da2};
        class_force(c, &kd);
      }
    }
    if (!coef.empty()) {
      const StepCoef& s = coef.back();
      cola_kick_drift(c.field(0), c.field(0), c.field(1), c.field(4), c.field(2), c.field(3), c.np,
                      c.np, m15om, s.kf, s.q1, s.q2, s.dt, s.da1, s.da2);
    }
    FS2_HIP(hipEventRecord(b->join[k], c.s));
  }
  for (size_t k = 0; k < b->cls.size(); ++k) FS2_HIP(hipStreamWaitEvent(main, b->join[k], 0));
}

void restore_ic(fs2_batch* b) {
  for (BatchClass& c : b->cls)
    FS2_HIP(hipMemcpyAsync(c.planes, c.ic, 4 * 3 * (size_t)c.np * sizeof(float),
                           hipMemcpyDeviceToDevice, library_stream()));
}

}  // namespace
}  // namespace fs2

using namespace fs2;

extern "C" {

int fs2_batch_create(int nclass, const int* nc_lpt, const int* nc_pm, const int* nbox,
                     const double* boxsize, fs2_batch** out) {
  return guard([&] {
    if (!out) fail(FS2_ERR_VALUE, "null output handle");
    *out = nullptr;
    if (nclass < 1 || !nc_lpt || !nc_pm || !nbox || !boxsize)
      fail(FS2_ERR_VALUE, "need at least one class and its sizes");
    fs2_batch* b = new fs2_batch();
    try {
      for (int k = 0; k < nclass; ++k) {
        const int nl = nc_lpt[k], nm = nc_pm[k];
        if (nl < 4 || nl % 2 != 0 || !fft_supported(nl))
          fail(FS2_ERR_VALUE, "class %d: nc_lpt = %d not supported for 2LPT", k, nl);
        if (nm < 4 || !fft_supported(nm))
          fail(FS2_ERR_VALUE, "class %d: nc_pm = %d is not a supported FFT size", k, nm);
        fft_prepare(nl);   // twiddle tables now, never inside a graph capture
        fft_prepare(nm);
        if (nbox[k] < 1) fail(FS2_ERR_VALUE, "class %d: nbox must be >= 1", k);
        if (!(boxsize[k] > 0)) fail(FS2_ERR_VALUE, "class %d: boxsize must be positive", k);
        BatchClass c;
        c.nl = nl;
        c.nm = nm;
        c.nbox = nbox[k];
        c.boxsize = boxsize[k];
        c.boxsize_f = (float)boxsize[k];
        const double pm_factor = (double)nm / nl;
        c.fac = (float)(pm_factor * pm_factor * pm_factor);
        c.f1 = (float)(-1.0 / std::pow((double)nm, 3.0) / (2.0 * M_PI / (double)c.boxsize_f));
        c.npb = (int64_t)nl * nl * nl;
        c.np = c.npb * c.nbox;
        if (c.np > (int64_t)0x7fffffff) fail(FS2_ERR_VALUE, "class %d: too many particles", k);
        c.tile.nbox = c.nbox;
        b->cls.push_back(c);
        BatchClass& cc = b->cls.back();
        cc.planes = (float*)dmalloc(5 * 3 * (size_t)cc.np * sizeof(float));
        cc.ic = (float*)dmalloc(4 * 3 * (size_t)cc.np * sizeof(float));
        const size_t mbytes = grid_floats(nm) * sizeof(float) * cc.nbox;
        cc.mesh = (float*)dmalloc(mbytes);
        for (auto& g : cc.F) g = (float*)dmalloc(mbytes);
        ensure_paint_scratch(cc.paint, (size_t)cc.np, max_bins(cc));
        FS2_HIP(hipStreamCreateWithFlags(&cc.s, hipStreamNonBlocking));
        hipEvent_t e;
        FS2_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
        b->join.push_back(e);
      }
      FS2_HIP(hipEventCreateWithFlags(&b->fork, hipEventDisableTiming));
    } catch (...) {
      destroy(b);
      throw;
    }
    *out = b;
  });
}

void fs2_batch_destroy(fs2_batch* b) { destroy(b); }

int fs2_batch_init_lpt(fs2_batch* b, const fs2_power* ps, double a, const uint64_t* seeds) {
  return guard([&] {
    if (!b || !ps || !seeds) fail(FS2_ERR_VALUE, "null argument");
    (void)host::omega_m();
    const float D1 = (float)host::D_growth(a);             // lpt.cpp:156-175, 'cola' kind
    const float D2 = (float)host::D2_growth(a, (double)D1);
    size_t k = 0;
    for (BatchClass& c : b->cls) {
      const std::vector<double> amp = lpt_amp_table(c.nl, c.boxsize, ps);
      const size_t gbytes = grid_floats(c.nl) * sizeof(float);
      std::vector<float*> g;
      double* d_amp = nullptr;
      try {
        for (int i = 0; i < 7; ++i) g.push_back((float*)dmalloc(gbytes));
        d_amp = (double*)dmalloc(amp.size() * sizeof(double));
        FS2_HIP(hipMemcpyAsync(d_amp, amp.data(), amp.size() * sizeof(double),
                               hipMemcpyHostToDevice, stream()));
        float* G[6] = {g[1], g[2], g[3], g[4], g[5], g[6]};
        for (int box = 0; box < c.nbox; ++box, ++k) {
          white_noise(g[0], c.nl, seeds[k]);
          lpt_displacements(g[0], G, c.nl, c.boxsize, d_amp);
          const int64_t off = (int64_t)box * c.npb;
          LatticeArgs la{c.field(0) + off, c.field(1) + off, c.field(2) + off, c.field(3) + off,
                         nullptr, c.np};
          const float* psi1[3] = {G[1], G[2], g[0]};
          const float* psi2[3] = {G[3], G[4], G[5]};
          lpt_lattice(la, psi1, psi2, c.nl, (float)(c.boxsize / c.nl), lpt_offset_value(), D1,
                      D2, 0.f, 0.f, 1.0 / std::pow((double)c.nl, 3.0));
        }
        FS2_HIP(hipMemcpyAsync(c.ic, c.planes, 4 * 3 * (size_t)c.np * sizeof(float),
                               hipMemcpyDeviceToDevice, stream()));
        sync();
      } catch (...) {
        for (float* p : g) dfree(p);
        dfree(d_amp);
        throw;
      }
      for (float* p : g) dfree(p);
      dfree(d_amp);
    }
    b->have_ic = true;
    b->a_ic = a;
  });
}

int fs2_batch_set_tile(fs2_batch* b, int cls, int tx, int ty, int tz) {
  return guard([&] {
    if (!b || cls < 0 || cls >= (int)b->cls.size()) fail(FS2_ERR_VALUE, "bad class index");
    BatchClass& c = b->cls[cls];
    const int t[3] = {tx, ty, tz};
    if (!shape_ok(c.nm, t))
      fail(FS2_ERR_VALUE, "tile %dx%dx%d not usable for the %d^3 mesh", tx, ty, tz, c.nm);
    TileOpts o = c.tile;
    o.tx = tx;
    o.ty = ty;
    o.tz = tz;
    ensure_paint_scratch(c.paint, (size_t)c.np, std::max(c.paint.nbins, paint_tiled_bins(c.nm, o)));
    c.tile = o;
    drop_graph(b);   // the captured launches carry the old shape
  });
}

int fs2_batch_get_tile(const fs2_batch* b, int cls, int* tx, int* ty, int* tz) {
  return guard([&] {
    if (!b || cls < 0 || cls >= (int)b->cls.size()) fail(FS2_ERR_VALUE, "bad class index");
    const BatchClass& c = b->cls[cls];
    int ext[3];
    tile_extents(c.nm, c.tile, ext);
    *tx = ext[0];
    *ty = ext[1];
    *tz = ext[2];
  });
}

int fs2_batch_autotune(fs2_batch* b, int reps, double* best_ms) {
  return guard([&] {
    if (!b) fail(FS2_ERR_VALUE, "null handle");
    if (!b->have_ic) fail(FS2_ERR_RUNTIME, "autotune needs initial conditions (fs2_batch_init_lpt)");
    if (reps < 1) reps = 1;
    const bool prof = profiling_enabled();
    set_profiling(false);
    hipEvent_t e0, e1;
    FS2_HIP(hipEventCreate(&e0));
    FS2_HIP(hipEventCreate(&e1));
    try {
      for (size_t k = 0; k < b->cls.size(); ++k) {
        BatchClass& c = b->cls[k];
        StreamScope scope(library_stream());
        restore_ic(b);
        // force meshes once (their values do not change the timing)
        class_force(c);
        const PosArgs pos{c.field(0), c.np, c.np};
        double best = 1e300;
        TileOpts best_o = c.tile;
        // meshes that none of the listed shapes divides (e.g. 20^3, 30^3)
        // time their current tile alone
        std::vector<std::array<int, 3>> shapes;
        for (const auto& t : kShapes)
          if (shape_ok(c.nm, t)) shapes.push_back({t[0], t[1], t[2]});
        if (shapes.empty()) {
          int ext[3];
          tile_extents(c.nm, c.tile, ext);
          shapes.push_back({ext[0], ext[1], ext[2]});
        }
        for (const auto& t : shapes) {
          TileOpts o = c.tile;
          o.tx = t[0];
          o.ty = t[1];
          o.tz = t[2];
          auto run = [&] {
            paint_tiled(pos, c.mesh, c.nm, c.boxsize_f, c.fac, true, true, c.paint, o);
            readout_tiled(c.F[0], c.F[1], c.F[2], c.field(4), c.np, c.np, c.nm, c.boxsize_f,
                          c.paint, o);
          };
          run();   // warm-up
          FS2_HIP(hipEventRecord(e0, stream()));
          for (int r = 0; r < reps; ++r) run();
          FS2_HIP(hipEventRecord(e1, stream()));
          FS2_HIP(hipEventSynchronize(e1));
          float ms = 0.f;
          FS2_HIP(hipEventElapsedTime(&ms, e0, e1));
          const double per = ms / reps;
          logf(1, "fs2_batch_autotune: class %zu tile %dx%dx%d %.4f ms\n", k, t[0], t[1], t[2],
               per);
          if (per < best) {
            best = per;
            best_o = o;
          }
        }
        c.tile = best_o;
        if (best_ms) best_ms[k] = best;
      }
      restore_ic(b);
      sync();
    } catch (...) {
      (void)hipEventDestroy(e0);
      (void)hipEventDestroy(e1);
      set_profiling(prof);
      throw;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    set_profiling(prof);
    drop_graph(b);
  });
}

int fs2_batch_set_schedule(fs2_batch* b, double a0, double a1, int nstep) {
  return guard([&] {
    if (!b) fail(FS2_ERR_VALUE, "null handle");
    if (nstep < 1 || !(a0 > 0) || !(a1 > a0)) fail(FS2_ERR_VALUE, "need nstep >= 1, 0 < a0 < a1");
    b->a0 = a0;
    b->a1 = a1;
    b->nstep = nstep;
    b->have_schedule = true;
    drop_graph(b);
  });
}

int fs2_batch_reset(fs2_batch* b) {
  return guard([&] {
    if (!b || !b->have_ic) fail(FS2_ERR_RUNTIME, "no initial conditions to restore");
    restore_ic(b);
    sync();
  });
}

int fs2_batch_capture(fs2_batch* b) {
  return guard([&] {
    if (!b || !b->have_schedule) fail(FS2_ERR_RUNTIME, "set a schedule before capturing");
    drop_graph(b);
    // twiddle tables and the like are allocated on first use: one eager
    // force pass of every class first, so capture records kernels only
    for (BatchClass& c : b->cls) {
      StreamScope scope(c.s);
      class_force(c);
      FS2_HIP(hipStreamSynchronize(c.s));
    }
    const bool prof = profiling_enabled();
    set_profiling(false);
    hipStream_t main = library_stream();
    FS2_HIP(hipStreamBeginCapture(main, hipStreamCaptureModeThreadLocal));
    try {
      run_schedule(b);
    } catch (...) {
      hipGraph_t g = nullptr;
      (void)hipStreamEndCapture(main, &g);
      if (g) (void)hipGraphDestroy(g);
      set_profiling(prof);
      throw;
    }
    set_profiling(prof);
    FS2_HIP(hipStreamEndCapture(main, &b->graph));
    FS2_HIP(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
  });
}

int fs2_batch_run(fs2_batch* b, int use_graph) {
  return guard([&] {
    if (!b || !b->have_schedule) fail(FS2_ERR_RUNTIME, "set a schedule before running");
    if (use_graph) {
      if (!b->exec) {
        int rc = fs2_batch_capture(b);
        if (rc) fail(rc, "%s", fs2_last_error());
      }
      FS2_HIP(hipGraphLaunch(b->exec, library_stream()));
    } else {
      run_schedule(b);
    }
  });
}

int fs2_batch_get_field(fs2_batch* b, int cls, int box, int field, float* out) {
  return guard([&] {
    if (!b || cls < 0 || cls >= (int)b->cls.size()) fail(FS2_ERR_VALUE, "bad class index");
    BatchClass& c = b->cls[cls];
    if (box < 0 || box >= c.nbox) fail(FS2_ERR_VALUE, "bad box index");
    if (field != FS2_FIELD_X && field != FS2_FIELD_V && field != FS2_FIELD_DX1 &&
        field != FS2_FIELD_DX2 && field != FS2_FIELD_FORCE)
      fail(FS2_ERR_VALUE, "bad field %d", field);
    // planes are x, v, dx1, dx2, force
    const int k = field == FS2_FIELD_X ? 0 : field == FS2_FIELD_V ? 1 : field == FS2_FIELD_DX1 ? 2
                : field == FS2_FIELD_DX2 ? 3 : 4;
    sync();
    std::vector<float> tmp(3 * (size_t)c.npb);
    for (int a = 0; a < 3; ++a)
      FS2_HIP(hipMemcpy(tmp.data() + a * c.npb, c.field(k) + a * c.np + box * c.npb,
                        c.npb * sizeof(float), hipMemcpyDeviceToHost));
    for (int64_t i = 0; i < c.npb; ++i)
      for (int a = 0; a < 3; ++a) out[3 * i + a] = tmp[a * c.npb + i];
  });
}

int fs2_batch_info(const fs2_batch* b, int* nclass, uint64_t* np_total) {
  return guard([&] {
    if (!b) fail(FS2_ERR_VALUE, "null handle");
    if (nclass) *nclass = (int)b->cls.size();
    uint64_t n = 0;
    for (const BatchClass& c : b->cls) n += (uint64_t)c.np;
    if (np_total) *np_total = n;
  });
}

}  // extern "C"
