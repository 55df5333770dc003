// Differentiable COLA run: forward with per-step checkpoints and the
// reverse-mode sweep (vector-Jacobian product) back to the white noise.
//
// No reference counterpart: fs2 has no adjoint.  The forward is the
// reference's 'cola' pipeline -- 2LPT (py/lpt.cpp:100-225, 439-591) with
// v0 = 0, then nstep kick/drift steps (py/cola.cpp:48-127) on the step
// schedule of script/cola.py:37-44, forces by the PM chain
// (py/pm.cpp:57-508), and the final CIC density.
//
// Checkpoints.  Positions of every step are always kept (12 bytes x
// particles x (nstep + 1)).  With FS2_DIFF_STORE the forward also keeps each
// step's three force meshes and its particle bins, so the backward sweep
// recomputes nothing; with FS2_DIFF_RECOMPUTE the backward re-bins and
// re-transforms the density of each step from its positions.  The default
// (FS2_DIFF_AUTO) stores when that fits in half of the free device memory,
// which on a 288 GB part covers every configuration up to 512^3 / 1024^3.
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
  bool store = false;         // force meshes and bins of every step kept
  int variant = FS2_PAINT_SORTED;   // FS2_PAINT_GLOBAL: atomic paint + gather kernels
  bool have_forward = false;
  int shared_bins_step = -1;  // step whose positions the shared bins hold
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
  float* xs = nullptr;        // positions [nstep + 1][3][np]
  float* xbar = nullptr;
  float* vbar = nullptr;
  float* dx1bar = nullptr;
  float* dx2bar = nullptr;
  float* fbar = nullptr;
  float* mesh = nullptr;      // nm grids
  float* F[3] = {};           // force meshes (recompute mode)
  float* P[3] = {};
  float* Fs = nullptr;        // store mode: [nstep][3] force meshes
  fs2::TileOpts tile;         // paint / readout tile extents
  fs2::PaintScratch paint;    // shared bins
  std::vector<fs2::PaintScratch> bins;   // store mode: per-step rec / offsets

  float* x_at(int i) const { return xs + (size_t)i * 3 * np; }
  float* force_mesh(int i, int a) const {
    return store ? Fs + ((size_t)i * 3 + a) * fs2::grid_floats(nm) : F[a];
  }
  const fs2::PaintScratch& bins_of(int i) const {
    return store && i < nstep ? bins[i] : paint;
  }
};

namespace fs2 {
namespace {

void destroy(fs2_diff* d) {
  if (!d) return;
  (void)hipStreamSynchronize(stream());
  for (void* p : {(void*)d->amp, (void*)d->noise, (void*)d->S, (void*)d->v, (void*)d->dx1,
                  (void*)d->dx2, (void*)d->f, (void*)d->xs, (void*)d->xbar, (void*)d->vbar,
                  (void*)d->dx1bar, (void*)d->dx2bar, (void*)d->fbar, (void*)d->mesh,
                  (void*)d->Fs})
    dfree(p);
  for (float* g : d->G) dfree(g);
  for (float* g : d->H) dfree(g);
  for (float* g : d->F) dfree(g);
  for (float* g : d->P) dfree(g);
  for (PaintScratch& b : d->bins) {
    dfree(b.rec);
    dfree(b.offsets);
  }
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

bool tiled(const fs2_diff* d) { return d->variant == FS2_PAINT_SORTED; }

// CIC density fac * sum W - 1 of the positions of step i into d->mesh,
// binning them into bins (tiled variant); wrap = the positions are wrapped
// into the box in place first.
void density(fs2_diff* d, int i, const PaintScratch& bins, bool wrap) {
  const PosArgs pos{d->x_at(i), d->np, d->np};
  if (tiled(d)) {
    paint_tiled(pos, d->mesh, d->nm, d->boxsize_f, d->fac, wrap, true, bins, d->tile);
    if (&bins == &d->paint) d->shared_bins_step = i;
  } else {
    fill_grid(d->mesh, d->nm, -1.0f);
    paint_atomic(pos, d->mesh, d->nm, d->boxsize_f, d->fac, wrap);
  }
}

void grad_readout(fs2_diff* d, int i, const float* rhobar, const float* const F[3],
                  const float* fbar) {
  if (tiled(d))
    grad_readout_tiled(rhobar, d->fac, F, fbar, d->xbar, d->np, d->np, d->nm, d->boxsize_f,
                       d->bins_of(i), d->tile);
  else
    adj_grad_readout(d->x_at(i), d->np, d->np, rhobar, d->fac, F, fbar, d->xbar, d->nm,
                     d->boxsize_f);
}

}  // namespace
}  // namespace fs2

using namespace fs2;

extern "C" {

int fs2_diff_create(int nc_lpt, int nc_pm, double boxsize, double a0, double a1, int nstep,
                    const fs2_power* ps, int checkpoint, fs2_diff** out) {
  return guard([&] {
    if (!out) fail(FS2_ERR_VALUE, "null output handle");
    *out = nullptr;
    if (!ps) fail(FS2_ERR_VALUE, "null power spectrum");
    if (nc_lpt < 4 || nc_lpt % 4 != 0 || !fft_supported(nc_lpt))
      fail(FS2_ERR_VALUE, "nc_lpt = %d not supported for 2LPT", nc_lpt);
    if (nc_pm < 4 || !fft_supported(nc_pm))
      fail(FS2_ERR_VALUE, "nc_pm = %d is not a supported FFT size", nc_pm);
    if (boxsize <= 0) fail(FS2_ERR_VALUE, "boxsize must be positive");
    if (nstep < 0) fail(FS2_ERR_VALUE, "nstep must be >= 0");
    if (!(a0 > 0.0) || !(a1 >= a0)) fail(FS2_ERR_VALUE, "need 0 < a0 <= a1");
    if (checkpoint != FS2_DIFF_AUTO && checkpoint != FS2_DIFF_RECOMPUTE &&
        checkpoint != FS2_DIFF_STORE)
      fail(FS2_ERR_VALUE, "unknown checkpoint policy %d", checkpoint);
    const double om = host::omega_m();   // cosmology must be initialised
    (void)paint_tiled_bins(nc_pm);       // tile extents are powers of two for every FFT size

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
        double kf, q1, q2, dt, da1, da2;
        int rc = fs2_cola_kick_coefficients(a_v, a_x, a_vel, &kf, &q1, &q2);
        if (rc) fail(rc, "%s", fs2_last_error());
        a_v = a_vel;
        rc = fs2_cola_drift_coefficients(a_x, a_v, a_pos, &dt, &da1, &da2);
        if (rc) fail(rc, "%s", fs2_last_error());
        a_x = a_pos;
        fs2_diff::Step s;
        s.kf = (float)kf;
        s.q1 = (float)q1;
        s.q2 = (float)q2;
        s.dt = dt;
        s.da1 = (float)da1;
        s.da2 = (float)da2;
        s.c = (float)(-1.5 * om * (double)s.kf);
        d->steps.push_back(s);
      }

      const size_t lbytes = grid_floats(nc_lpt) * sizeof(float);
      const size_t mbytes = grid_floats(nc_pm) * sizeof(float);
      const size_t pbytes = 3 * (size_t)d->np * sizeof(float);
      const size_t nbins = paint_tiled_bins(nc_pm, d->tile);
      const size_t store_bytes =
          (size_t)nstep * (3 * mbytes + (size_t)d->np * sizeof(float4) + (nbins + 1) * 4);
      if (checkpoint == FS2_DIFF_AUTO) {
        size_t free_b = 0, total_b = 0;
        FS2_HIP(hipMemGetInfo(&free_b, &total_b));
        const size_t base = 10 * lbytes + 10 * pbytes + pbytes * (size_t)(nstep + 1) +
                            7 * mbytes + (size_t)d->np * 20;
        d->store = base + store_bytes < free_b / 2;
      } else {
        d->store = checkpoint == FS2_DIFF_STORE;
      }

      const std::vector<double> amp = lpt_amp_table(nc_lpt, boxsize, ps);
      d->amp = (double*)dmalloc(amp.size() * sizeof(double));
      FS2_HIP(hipMemcpyAsync(d->amp, amp.data(), amp.size() * sizeof(double),
                             hipMemcpyHostToDevice, stream()));
      d->noise = (float*)dmalloc(lbytes);
      d->S = (float*)dmalloc(lbytes);
      for (auto& g : d->G) g = (float*)dmalloc(lbytes);
      for (auto& g : d->H) g = (float*)dmalloc(lbytes);
      for (float** pp : {&d->v, &d->dx1, &d->dx2, &d->f, &d->xbar, &d->vbar, &d->dx1bar,
                         &d->dx2bar, &d->fbar})
        *pp = (float*)dmalloc(pbytes);
      d->xs = (float*)dmalloc(pbytes * (size_t)(nstep + 1));
      d->mesh = (float*)dmalloc(mbytes);
      for (auto& g : d->P) g = (float*)dmalloc(mbytes);
      ensure_paint_scratch(d->paint, (size_t)d->np, nbins);
      if (d->store) {
        if (nstep > 0) d->Fs = (float*)dmalloc(3 * mbytes * (size_t)nstep);
        for (int i = 0; i < nstep; ++i) {
          PaintScratch b = d->paint;   // shares keys / counts / cursor / scan sums
          b.rec = nullptr;
          b.offsets = nullptr;
          d->bins.push_back(b);
          d->bins.back().rec = (float4*)dmalloc((size_t)d->np * sizeof(float4));
          d->bins.back().offsets = (unsigned*)dmalloc((nbins + 1) * sizeof(unsigned));
        }
      } else {
        for (auto& g : d->F) g = (float*)dmalloc(mbytes);
      }
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

int fs2_diff_params(const fs2_diff* d, int* nc_lpt, int* nc_pm, int* nstep, uint64_t* np,
                    int* stores_forces) {
  return guard([&] {
    if (!d) fail(FS2_ERR_VALUE, "null handle");
    if (nc_lpt) *nc_lpt = d->nl;
    if (nc_pm) *nc_pm = d->nm;
    if (nstep) *nstep = d->nstep;
    if (np) *np = (uint64_t)d->np;
    if (stores_forces) *stores_forces = d->store ? 1 : 0;
  });
}

int fs2_diff_set_paint_variant(fs2_diff* d, int variant) {
  return guard([&] {
    if (!d) fail(FS2_ERR_VALUE, "null handle");
    if (variant != FS2_PAINT_SORTED && variant != FS2_PAINT_GLOBAL)
      fail(FS2_ERR_VALUE, "differentiable run supports the sorted and global paint variants");
    d->variant = variant;
    d->have_forward = false;   // checkpoints of the other variant are not reused
  });
}

int fs2_diff_forward(fs2_diff* d, const float* noise, float* delta) {
  return guard([&] {
    if (!d) fail(FS2_ERR_VALUE, "null handle");
    if (!noise || !delta) fail(FS2_ERR_VALUE, "null device buffer");
    const int nl = d->nl, nm = d->nm;
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
      float* F[3] = {d->force_mesh(i, 0), d->force_mesh(i, 1), d->force_mesh(i, 2)};
      density(d, i, d->bins_of(i), true);   // the checkpoint is the wrapped position
      fft_force_meshes(d->mesh, F, nm, d->f1);
      if (tiled(d))
        readout_tiled(F[0], F[1], F[2], d->f, np, np, nm, d->boxsize_f, d->bins_of(i), d->tile);
      else
        readout3(PosArgs{d->x_at(i), np, np}, F[0], F[1], F[2], d->f, nm, d->boxsize_f);
      // kick, then drift from checkpoint i into checkpoint i + 1
      cola_kick_drift(d->x_at(i + 1), d->x_at(i), d->v, d->f, d->dx1, d->dx2, np, np, m15om,
                      s.kf, s.q1, s.q2, s.dt, s.da1, s.da2);
    }
    density(d, d->nstep, d->paint, true);
    from_grid(delta, d->mesh, nm);
    sync();
    d->have_forward = true;
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
    if (tiled(d) && d->shared_bins_step != d->nstep) density(d, d->nstep, d->paint, false);
    to_grid(d->P[0], delta_bar, nm);
    grad_readout(d, d->nstep, d->P[0], nullptr, nullptr);

    for (int i = d->nstep - 1; i >= 0; --i) {
      const fs2_diff::Step& s = d->steps[i];
      adj_cola_step(d->xbar, d->vbar, d->dx1bar, d->dx2bar, d->fbar, n3, (float)s.dt, s.da1,
                    s.da2, s.c, s.q1, s.q2);
      float* F[3] = {d->force_mesh(i, 0), d->force_mesh(i, 1), d->force_mesh(i, 2)};
      if (!d->store) {   // force meshes at x_i, recomputed from the checkpoint
        density(d, i, d->paint, false);
        fft_force_meshes(d->mesh, F, nm, d->f1);
      }
      // rhobar = M^T (paint of fbar) = -sum_a M_a P_a
      if (tiled(d)) {
        paint_values3_tiled(d->fbar, np, np, d->P, nm, d->boxsize_f, d->bins_of(i), d->tile);
      } else {
        for (float* g : d->P) fill_grid(g, nm, 0.0f);
        adj_paint_values3(d->x_at(i), d->fbar, np, np, d->P, nm, d->boxsize_f);
      }
      fft_adjoint_green(d->P, nm, d->f1);
      grad_readout(d, i, d->P[0], F, d->fbar);
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
