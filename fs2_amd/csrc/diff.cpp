// Differentiable COLA run: forward with per-step position checkpoints and
// the reverse-mode sweep (vector-Jacobian product) back to the white noise.
//
// No reference counterpart: fs2 has no adjoint.  The forward is the
// reference's 'cola' pipeline -- 2LPT (py/lpt.cpp:100-225, 439-591) with
// v0 = 0, then nstep kick/drift steps (py/cola.cpp:48-127) on the step
// schedule of script/cola.py:37-44, forces by the PM chain
// (py/pm.cpp:57-508), and the final CIC density.
//
// Memory is traded for time the MI355X way: positions of every step are
// kept in HBM (12 bytes x particles x (nstep + 1)), so the backward sweep
// only recomputes one density + force-mesh transform per step and never
// re-runs the forward from the start.
#include <cmath>
#include <vector>

#include "fs2_internal.h"

struct fs2_diff {
  int nl = 0, nm = 0, nstep = 0;
  double boxsize = 0.0, a0 = 0.0, a1 = 0.0;
  float boxsize_f = 0.f, fac = 1.f, f1 = 0.f;
  float D1 = 0.f, D2 = 0.f;   // lattice growth factors at a0 (FP32, as the lattice kernel)
  struct Step {
    float kf, q1, q2, da1, da2;
    double dt;
    float c;                  // d v / d (kick sum) = -1.5 om kf
  };
  std::vector<Step> steps;
  int64_t np = 0;
  bool tiled = true;          // tiled deterministic paint (else global atomics)
  bool have_forward = false;
  uint64_t forward_count = 0;
  // device buffers
  double* amp = nullptr;
  float* noise = nullptr;     // padded copy of the last forward's noise (nl grid)
  float* S = nullptr;
  float* G[6] = {};
  float* H[3] = {};
  float* v = nullptr;         // planes [3][np]
  float* dx1 = nullptr;
  float* dx2 = nullptr;
  float* f = nullptr;
  float* xs = nullptr;        // checkpoints [nstep + 1][3][np]
  float* xbar = nullptr;
  float* vbar = nullptr;
  float* dx1bar = nullptr;
  float* dx2bar = nullptr;
  float* fbar = nullptr;
  float* mesh = nullptr;      // nm grids
  float* F[3] = {};
  float* P[3] = {};
  fs2::PaintScratch paint;

  float* x_at(int i) const { return xs + (size_t)i * 3 * np; }
};

namespace fs2 {
namespace {

std::vector<void*> buffers_of(fs2_diff* d) {
  std::vector<void*> b = {d->amp, d->noise, d->S, d->v, d->dx1, d->dx2, d->f, d->xs, d->xbar,
                          d->vbar, d->dx1bar, d->dx2bar, d->fbar, d->mesh};
  for (float* g : d->G) b.push_back(g);
  for (float* g : d->H) b.push_back(g);
  for (float* g : d->F) b.push_back(g);
  for (float* g : d->P) b.push_back(g);
  return b;
}

void destroy(fs2_diff* d) {
  if (!d) return;
  (void)hipStreamSynchronize(stream());
  for (void* p : buffers_of(d)) dfree(p);
  free_paint_scratch(d->paint);
  delete d;
}

// contiguous n^3 <-> padded grid, device to device
void to_grid(float* grid, const float* src, int n) {
  FS2_HIP(hipMemcpy2DAsync(grid, grid_rpitch(n) * sizeof(float), src, n * sizeof(float),
                           n * sizeof(float), (size_t)n * n, hipMemcpyDeviceToDevice, stream()));
}
void from_grid(float* dst, const float* grid, int n) {
  FS2_HIP(hipMemcpy2DAsync(dst, n * sizeof(float), grid, grid_rpitch(n) * sizeof(float),
                           n * sizeof(float), (size_t)n * n, hipMemcpyDeviceToDevice, stream()));
}

// CIC density fac * sum W - 1 of the positions at x into d->mesh; wraps x in place.
void density(fs2_diff* d, float* x, bool wrap) {
  const PosArgs pos{x, d->np, d->np};
  if (d->tiled) {
    paint_tiled(pos, d->mesh, d->nm, d->boxsize_f, d->fac, wrap, true, d->paint);
  } else {
    fill_grid(d->mesh, d->nm, -1.0f);
    paint_atomic(pos, d->mesh, d->nm, d->boxsize_f, d->fac, wrap);
  }
}

}  // namespace
}  // namespace fs2

using namespace fs2;

extern "C" {

int fs2_diff_create(int nc_lpt, int nc_pm, double boxsize, double a0, double a1, int nstep,
                    const fs2_power* ps, fs2_diff** out) {
  return guard([&] {
    if (!out) fail(FS2_ERR_VALUE, "null output handle");
    *out = nullptr;
    if (!ps) fail(FS2_ERR_VALUE, "null power spectrum");
    if (nc_lpt < 4 || nc_lpt % 4 != 0 || !fft_supported(nc_lpt))
      fail(FS2_ERR_VALUE, "nc_lpt = %d not supported for 2LPT", nc_lpt);
    if (nc_pm < 2 || !fft_supported(nc_pm))
      fail(FS2_ERR_VALUE, "nc_pm = %d is not a supported FFT size", nc_pm);
    if (boxsize <= 0) fail(FS2_ERR_VALUE, "boxsize must be positive");
    if (nstep < 0) fail(FS2_ERR_VALUE, "nstep must be >= 0");
    if (!(a0 > 0.0) || !(a1 >= a0)) fail(FS2_ERR_VALUE, "need 0 < a0 <= a1");
    const double om = host::omega_m();   // cosmology must be initialised

    fs2_diff* d = new fs2_diff();
    try {
      d->nl = nc_lpt;
      d->nm = nc_pm;
      d->nstep = nstep;
      d->boxsize = boxsize;
      d->boxsize_f = (float)boxsize;
      d->a0 = a0;
      d->a1 = a1;
      const double pm_factor = (double)nc_pm / nc_lpt;
      d->fac = (float)(pm_factor * pm_factor * pm_factor);
      d->f1 = (float)(-1.0 / std::pow((double)nc_pm, 3.0) / (2.0 * M_PI / (double)d->boxsize_f));
      d->D1 = (float)host::D_growth(a0);
      d->D2 = (float)host::D2_growth(a0, (double)d->D1);
      d->np = (int64_t)nc_lpt * nc_lpt * nc_lpt;

      // step coefficients of the forward, in the order fs2_cola_kick/drift use them
      double a_x = a0, a_v = a0;
      for (int i = 0; i < nstep; ++i) {
        const double a_vel = a0 + (a1 - a0) / nstep * (i + 0.5);
        const double a_pos = a0 + (a1 - a0) / nstep * (i + 1.0);
        fs2_diff::Step s;
        double kf, q1, q2, dt, da1, da2;
        int rc = fs2_cola_kick_coefficients(a_v, a_x, a_vel, &kf, &q1, &q2);
        if (rc) fail(rc, "%s", fs2_last_error());
        a_v = a_vel;
        rc = fs2_cola_drift_coefficients(a_x, a_v, a_pos, &dt, &da1, &da2);
        if (rc) fail(rc, "%s", fs2_last_error());
        a_x = a_pos;
        s.kf = (float)kf;
        s.q1 = (float)q1;
        s.q2 = (float)q2;
        s.dt = dt;
        s.da1 = (float)da1;
        s.da2 = (float)da2;
        s.c = (float)(-1.5 * om * (double)s.kf);
        d->steps.push_back(s);
      }

      try {
        (void)paint_tiled_bins(nc_pm);
      } catch (const Error&) {
        d->tiled = false;   // tile extents must be powers of two; global atomics otherwise
      }

      const std::vector<double> amp = lpt_amp_table(nc_lpt, boxsize, ps);
      d->amp = (double*)dmalloc(amp.size() * sizeof(double));
      FS2_HIP(hipMemcpyAsync(d->amp, amp.data(), amp.size() * sizeof(double),
                             hipMemcpyHostToDevice, stream()));
      const size_t lbytes = grid_floats(nc_lpt) * sizeof(float);
      const size_t mbytes = grid_floats(nc_pm) * sizeof(float);
      const size_t pbytes = 3 * (size_t)d->np * sizeof(float);
      d->noise = (float*)dmalloc(lbytes);
      d->S = (float*)dmalloc(lbytes);
      for (auto& g : d->G) g = (float*)dmalloc(lbytes);
      for (auto& g : d->H) g = (float*)dmalloc(lbytes);
      for (float** pp : {&d->v, &d->dx1, &d->dx2, &d->f, &d->xbar, &d->vbar, &d->dx1bar,
                         &d->dx2bar, &d->fbar})
        *pp = (float*)dmalloc(pbytes);
      d->xs = (float*)dmalloc(pbytes * (size_t)(nstep + 1));
      d->mesh = (float*)dmalloc(mbytes);
      for (auto& g : d->F) g = (float*)dmalloc(mbytes);
      for (auto& g : d->P) g = (float*)dmalloc(mbytes);
      if (d->tiled) ensure_paint_scratch(d->paint, (size_t)d->np, paint_tiled_bins(nc_pm));
      // padding columns of the LPT grids stay defined
      FS2_HIP(hipMemsetAsync(d->noise, 0, lbytes, stream()));
      sync();
    } catch (...) {
      destroy(d);
      throw;
    }
    *out = d;
  });
}

void fs2_diff_destroy(fs2_diff* d) { destroy(d); }

int fs2_diff_params(const fs2_diff* d, int* nc_lpt, int* nc_pm, int* nstep, uint64_t* np) {
  return guard([&] {
    if (!d) fail(FS2_ERR_VALUE, "null handle");
    if (nc_lpt) *nc_lpt = d->nl;
    if (nc_pm) *nc_pm = d->nm;
    if (nstep) *nstep = d->nstep;
    if (np) *np = (uint64_t)d->np;
  });
}

int fs2_diff_forward(fs2_diff* d, const float* noise, float* delta) {
  return guard([&] {
    if (!d) fail(FS2_ERR_VALUE, "null handle");
    if (!noise || !delta) fail(FS2_ERR_VALUE, "null device buffer");
    const int nl = d->nl;
    const int64_t np = d->np;
    d->have_forward = false;

    // 2LPT from the noise ('cola' kind: v0 = 0)
    to_grid(d->noise, noise, nl);
    FS2_HIP(hipMemcpyAsync(d->S, d->noise, grid_floats(nl) * sizeof(float),
                           hipMemcpyDeviceToDevice, stream()));
    lpt_displacements(d->S, d->G, nl, d->boxsize, d->amp);
    LatticeArgs la{d->x_at(0), d->v, d->dx1, d->dx2, nullptr, np};
    const float* psi1[3] = {d->G[1], d->G[2], d->S};
    const float* psi2[3] = {d->G[3], d->G[4], d->G[5]};
    lpt_lattice(la, psi1, psi2, nl, (float)(d->boxsize / nl), lpt_offset_value(), d->D1, d->D2,
                0.f, 0.f, 1.0 / std::pow((double)nl, 3.0));

    const double m15om = -1.5 * host::omega_m();
    for (int i = 0; i < d->nstep; ++i) {
      const fs2_diff::Step& s = d->steps[i];
      float* x = d->x_at(i);
      density(d, x, true);   // wraps x_i in place: the checkpoint is the wrapped position
      fft_force_meshes(d->mesh, d->F, d->nm, d->f1);
      if (d->tiled)
        readout_tiled(d->F[0], d->F[1], d->F[2], d->f, np, np, d->nm, d->boxsize_f, d->paint);
      else
        readout3(PosArgs{x, np, np}, d->F[0], d->F[1], d->F[2], d->f, d->nm, d->boxsize_f);
      cola_kick(d->v, d->f, d->dx1, d->dx2, np, np, m15om, s.kf, s.q1, s.q2);
      float* xn = d->x_at(i + 1);
      FS2_HIP(hipMemcpyAsync(xn, x, 3 * (size_t)np * sizeof(float), hipMemcpyDeviceToDevice,
                             stream()));
      cola_drift(xn, d->v, d->dx1, d->dx2, np, np, s.dt, s.da1, s.da2);
    }
    density(d, d->x_at(d->nstep), true);
    from_grid(delta, d->mesh, d->nm);
    sync();
    d->have_forward = true;
    ++d->forward_count;
  });
}

int fs2_diff_backward(fs2_diff* d, const float* delta_bar, float* noise_bar) {
  return guard([&] {
    if (!d) fail(FS2_ERR_VALUE, "null handle");
    if (!delta_bar || !noise_bar) fail(FS2_ERR_VALUE, "null device buffer");
    if (!d->have_forward) fail(FS2_ERR_RUNTIME, "backward needs a forward on this handle first");
    const int nl = d->nl, nm = d->nm;
    const int64_t np = d->np, n3 = 3 * np;
    const size_t pbytes = (size_t)n3 * sizeof(float);
    for (float* b : {d->xbar, d->vbar, d->dx1bar, d->dx2bar})
      FS2_HIP(hipMemsetAsync(b, 0, pbytes, stream()));

    // final density: xbar_N = fac grad-readout of delta_bar
    to_grid(d->P[0], delta_bar, nm);
    adj_grad_readout(d->x_at(d->nstep), np, np, d->P[0], d->fac, nullptr, nullptr, d->xbar, nm,
                     d->boxsize_f);

    for (int i = d->nstep - 1; i >= 0; --i) {
      const fs2_diff::Step& s = d->steps[i];
      float* x = d->x_at(i);
      adj_cola_step(d->xbar, d->vbar, d->dx1bar, d->dx2bar, d->fbar, n3, (float)s.dt, s.da1,
                    s.da2, s.c, s.q1, s.q2);
      // force meshes at x_i, recomputed from the checkpoint
      density(d, x, false);
      fft_force_meshes(d->mesh, d->F, nm, d->f1);
      // rhobar = M^T (paint of fbar) = -sum_a M_a P_a
      for (float* g : d->P) fill_grid(g, nm, 0.0f);
      adj_paint_values3(x, d->fbar, np, np, d->P, nm, d->boxsize_f);
      for (float* g : d->P) fft_r2c_3d(g, nm);
      adj_green_sum(d->P, nm, d->f1);
      fft_c2r_3d(d->P[0], nm);
      adj_grad_readout(x, np, np, d->P[0], d->fac, d->F, d->fbar, d->xbar, nm, d->boxsize_f);
    }

    // x_0 = q + D1 dx1 + D2 dx2 (v_0 = 0)
    adj_axpy(d->dx1bar, d->xbar, d->D1, n3);
    adj_axpy(d->dx2bar, d->xbar, d->D2, n3);

    // 2LPT transposed: dx2 = B div2 / n^3 with div2 quadratic in psi_ij = S_ij w,
    // dx1 = T w  ->  wbar = -T dx1bar + sum_ij S_ij psibar_ij
    FS2_HIP(hipMemcpyAsync(d->S, d->noise, grid_floats(nl) * sizeof(float),
                           hipMemcpyDeviceToDevice, stream()));
    lpt_psi_fields(d->S, d->G, nl, d->boxsize, d->amp);
    for (int a = 0; a < 3; ++a) {
      plane_to_grid(d->dx2bar + a * np, d->H[a], nl);
      fft_r2c_3d(d->H[a], nl);
    }
    adj_lpt_divbar_k(d->H, nl, d->boxsize);
    fft_c2r_3d(d->H[0], nl);
    adj_lpt_div2(d->H[0], d->G, nl);
    for (float* g : d->G) fft_r2c_3d(g, nl);
    for (int a = 0; a < 3; ++a) {
      plane_to_grid(d->dx1bar + a * np, d->H[a], nl);
      fft_r2c_3d(d->H[a], nl);
    }
    adj_lpt_noisebar_k(d->H, d->G, nl, d->boxsize, d->amp);
    fft_c2r_3d(d->H[0], nl);
    from_grid(noise_bar, d->H[0], nl);
    sync();
  });
}

}  // extern "C"
