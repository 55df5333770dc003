// cpu_step.cpp -- FP32 CPU restatement of the reference's PM + COLA step
// loops, timed as the CPU baseline beside the GPU (bench only).
//
// Test and measurement infrastructure: only bench.py's CPU-baseline leg and
// tests/ load this library; the product path never does.
//
// It keeps what the reference's CPU path does per step, on the same data
// layout:
//   - particles as the 56-byte AoS record of py/particle.h:12-18;
//   - clear_density, py/pm.cpp:414-426 (delta = -1 on the nc^3 cells);
//   - pm_assign_cic_density, py/pm.cpp:57-139, with the `omp atomic` adds of
//     grid_assign (pm.cpp:32-39);
//   - compute_delta_k's copy of the transform into delta_k, pm.cpp:431-458;
//   - compute_force_mesh, pm.cpp:460-508: fk = i f2 delta_k with
//     f2 = f1 k_axis / |k|^2 in FP32 (Float), one axis at a time;
//   - force_at_particle_locations, pm.cpp:142-220, one axis per pass;
//   - cola_kick and cola_drift, py/cola.cpp:48-127, with coefficients from
//     FP64 and the per-particle arithmetic in FP32.
// Every `#pragma omp parallel for` sits where the reference has one (the
// reference's own build passes no -fopenmp; one thread reproduces that).
// The r2c and c2r transforms are the caller's (FFTW in the reference;
// scipy's pocketfft here).  Meshes are unpadded C-order nc^3 FP32; the
// half-complex arrays are [kx][ky][kz <= nc/2] complex64.
#include <cmath>
#include <cstdint>
#include <cstring>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

struct Particle {        // py/particle.h:12-18
  float x[3];
  float v[3];
  float dx1[3];
  float dx2[3];
  uint64_t id;
};
static_assert(sizeof(Particle) == 56, "56-byte particle record");

struct Cell {
  int i0, i1;
  float w0, w1;
};

inline Cell cell(float X, int nc) {
  Cell c;
  c.i0 = (int)X;                 // (int) cast, pm.cpp:95-97
  c.w1 = X - (float)c.i0;
  c.w0 = 1.0f - c.w1;
  if (c.i0 >= nc) c.i0 = 0;      // pm.cpp:109-111
  c.i1 = c.i0 + 1;
  if (c.i1 >= nc) c.i1 -= nc;
  // more than a cell outside [0, L) (never after the periodic wrap): index
  // from the far end so that no access leaves the mesh
  c.i0 = ((c.i0 % nc) + nc) % nc;
  c.i1 = ((c.i1 % nc) + nc) % nc;
  return c;
}

inline void grid_assign(float* d, int64_t nc, int ix, int iy, int iz, float f) {
#ifdef _OPENMP
#pragma omp atomic
#endif
  d[((int64_t)ix * nc + iy) * nc + iz] += f;
}

}  // namespace

extern "C" {

int cpu_step_threads(int n) {
#ifdef _OPENMP
  if (n > 0) omp_set_num_threads(n);
  return omp_get_max_threads();
#else
  (void)n;
  return 1;
#endif
}

int cpu_step_record_size(void) { return (int)sizeof(Particle); }

// pm.cpp:414-426
void cpu_clear_density(float* d, int nc) {
  const int64_t n = nc;
#ifdef _OPENMP
#pragma omp parallel for
#endif
  for (int64_t ix = 0; ix < n; ix++)
    for (int64_t iy = 0; iy < n; iy++)
      for (int64_t iz = 0; iz < n; iz++) d[(ix * n + iy) * n + iz] = -1.0f;
}

// util.h:14-24 periodic_wrapup_p, as pm_domain_send_positions applies it
void cpu_wrap(void* particles, int64_t np, float boxsize) {
  Particle* p = (Particle*)particles;
#ifdef _OPENMP
#pragma omp parallel for
#endif
  for (int64_t i = 0; i < np; i++)
    for (int k = 0; k < 3; k++) {
      if (p[i].x[k] < 0) p[i].x[k] += boxsize;
      if (p[i].x[k] >= boxsize) p[i].x[k] -= boxsize;
    }
}

// pm.cpp:57-139
void cpu_paint(const void* particles, int64_t np, float* d, int nc, float boxsize,
               double pm_factor) {
  const Particle* p = (const Particle*)particles;
  const float dx_inv = (float)nc / boxsize;
  const float fac = (float)(pm_factor * pm_factor * pm_factor);
#ifdef _OPENMP
#pragma omp parallel for
#endif
  for (int64_t i = 0; i < np; i++) {
    const Cell cx = cell(p[i].x[0] * dx_inv, nc);
    const Cell cy = cell(p[i].x[1] * dx_inv, nc);
    const Cell cz = cell(p[i].x[2] * dx_inv, nc);
    grid_assign(d, nc, cx.i0, cy.i0, cz.i0, fac * cx.w0 * cy.w0 * cz.w0);
    grid_assign(d, nc, cx.i0, cy.i0, cz.i1, fac * cx.w0 * cy.w0 * cz.w1);
    grid_assign(d, nc, cx.i0, cy.i1, cz.i0, fac * cx.w0 * cy.w1 * cz.w0);
    grid_assign(d, nc, cx.i0, cy.i1, cz.i1, fac * cx.w0 * cy.w1 * cz.w1);
    grid_assign(d, nc, cx.i1, cy.i0, cz.i0, fac * cx.w1 * cy.w0 * cz.w0);
    grid_assign(d, nc, cx.i1, cy.i0, cz.i1, fac * cx.w1 * cy.w0 * cz.w1);
    grid_assign(d, nc, cx.i1, cy.i1, cz.i0, fac * cx.w1 * cy.w1 * cz.w0);
    grid_assign(d, nc, cx.i1, cy.i1, cz.i1, fac * cx.w1 * cy.w1 * cz.w1);
  }
}

// pm.cpp:446-457: delta_k <- fk
void cpu_copy_delta_k(const float* fk, float* delta_k, int nc) {
  const int64_t n = nc, nz = nc / 2 + 1;
#ifdef _OPENMP
#pragma omp parallel for
#endif
  for (int64_t ix = 0; ix < n; ix++)
    std::memcpy(delta_k + 2 * ix * n * nz, fk + 2 * ix * n * nz, sizeof(float) * 2 * n * nz);
}

// pm.cpp:460-504: fk = i f2 delta_k, f2 = f1 k_axis / |k|^2, fk(0) = 0
void cpu_force_k(const float* delta_k, float* fk, int nc, float boxsize, int axis) {
  const int64_t n = nc, nz = nc / 2 + 1;
  const float f1 = (float)(-1.0 / std::pow((double)nc, 3.0) / (2.0 * M_PI / boxsize));
  fk[0] = 0.0f;
  fk[1] = 0.0f;
#ifdef _OPENMP
#pragma omp parallel for
#endif
  for (int64_t ix = 0; ix < n; ix++) {
    float k[3];
    k[0] = (float)(ix <= n / 2 ? ix : ix - n);
    for (int64_t iy = 0; iy < n; iy++) {
      k[1] = (float)(iy <= n / 2 ? iy : iy - n);
      const int64_t kzmin = (ix == 0 && iy == 0);
      for (int64_t iz = kzmin; iz < nz; iz++) {
        k[2] = (float)iz;
        const float f2 = f1 / (k[0] * k[0] + k[1] * k[1] + k[2] * k[2]) * k[axis];
        const int64_t idx = 2 * ((ix * n + iy) * nz + iz);
        fk[idx] = -f2 * delta_k[idx + 1];
        fk[idx + 1] = f2 * delta_k[idx];
      }
    }
  }
}

// pm.cpp:142-220: force[i][axis] from one real force mesh
void cpu_readout(const void* particles, int64_t np, const float* g, int nc, float boxsize,
                 int axis, float* force) {
  const Particle* p = (const Particle*)particles;
  const float dx_inv = (float)nc / boxsize;
  const int64_t n = nc;
#ifdef _OPENMP
#pragma omp parallel for
#endif
  for (int64_t i = 0; i < np; i++) {
    const Cell cx = cell(p[i].x[0] * dx_inv, nc);
    const Cell cy = cell(p[i].x[1] * dx_inv, nc);
    const Cell cz = cell(p[i].x[2] * dx_inv, nc);
    auto at = [&](int ix, int iy, int iz) { return g[((int64_t)ix * n + iy) * n + iz]; };
    float f = 0.0f;
    f += at(cx.i0, cy.i0, cz.i0) * cx.w0 * cy.w0 * cz.w0 +
         at(cx.i0, cy.i0, cz.i1) * cx.w0 * cy.w0 * cz.w1 +
         at(cx.i0, cy.i1, cz.i0) * cx.w0 * cy.w1 * cz.w0 +
         at(cx.i0, cy.i1, cz.i1) * cx.w0 * cy.w1 * cz.w1;
    f += at(cx.i1, cy.i0, cz.i0) * cx.w1 * cy.w0 * cz.w0 +
         at(cx.i1, cy.i0, cz.i1) * cx.w1 * cy.w0 * cz.w1 +
         at(cx.i1, cy.i1, cz.i0) * cx.w1 * cy.w1 * cz.w0 +
         at(cx.i1, cy.i1, cz.i1) * cx.w1 * cy.w1 * cz.w1;
    force[3 * i + axis] = f;
  }
}

// cola.cpp:77-86; kick_factor, q1, q2 arrive as the reference's Float values
void cpu_kick(void* particles, const float* force, int64_t np, double om, float kick_factor,
              float q1, float q2) {
  Particle* p = (Particle*)particles;
#ifdef _OPENMP
#pragma omp parallel for
#endif
  for (int64_t i = 0; i < np; i++) {
    const float* f = force + 3 * i;
    const float ax = -1.5 * om * (f[0] + p[i].dx1[0] * q1 + p[i].dx2[0] * q2);
    const float ay = -1.5 * om * (f[1] + p[i].dx1[1] * q1 + p[i].dx2[1] * q2);
    const float az = -1.5 * om * (f[2] + p[i].dx1[2] * q1 + p[i].dx2[2] * q2);
    p[i].v[0] += ax * kick_factor;
    p[i].v[1] += ay * kick_factor;
    p[i].v[2] += az * kick_factor;
  }
}

// cola.cpp:114-124; dt is a double as in the reference, da1, da2 Float
void cpu_drift(void* particles, int64_t np, double dt, float da1, float da2) {
  Particle* p = (Particle*)particles;
#ifdef _OPENMP
#pragma omp parallel for
#endif
  for (int64_t i = 0; i < np; i++)
    for (int k = 0; k < 3; k++)
      p[i].x[k] += p[i].v[k] * dt + (p[i].dx1[k] * da1 + p[i].dx2[k] * da2);
}

}  // extern "C"
