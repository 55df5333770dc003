// Adjoint (vector-Jacobian) kernels of the differentiable COLA run, gfx950.
//
// No reference counterpart (fs2 has no adjoint).  The forward operators are
// py/pm.cpp:57-220 (CIC paint/readout), pm.cpp:460-508 (Green x gradient),
// py/cola.cpp:48-127 (kick/drift) and py/lpt.cpp:439-591 (2LPT); their
// transposes here use:
//   * CIC:   d/dx of w0 = 1 - (X - i0), w1 = X - i0 with X = x n/L, so
//            dw0/dx = -n/L, dw1/dx = +n/L (piecewise linear, a.e.);
//   * Green: M_a = c2r . (i f2 k_a) . r2c is a convolution with an odd real
//            kernel, so M_a^T = -M_a (checked on the oracle, Nyquist
//            planes included);
//   * LPT:   the w -> dx1 maps are odd (T^T = -T), w -> psi_ij even
//            (S^T = S), div2 -> dx2 odd (B^T = -B).
#include <cmath>

#include "fs2_internal.h"

// CIC weights must round exactly as in the paint kernels.
#pragma clang fp contract(off)

namespace fs2 {
namespace {

constexpr int kBlock = 256;
inline unsigned blocks_for(int64_t n) { return (unsigned)((n + kBlock - 1) / kBlock); }

struct Cic1 {
  int i0, i1;
  float w0, w1;
};
__device__ __forceinline__ Cic1 cic1(float x, float dx_inv, int n) {
  Cic1 c;
  const float X = x * dx_inv;
  int i0 = (int)X;
  c.w1 = X - (float)i0;
  c.w0 = 1.f - c.w1;
  if (i0 >= n) i0 = 0;
  if ((unsigned)i0 >= (unsigned)n) i0 = ((i0 % n) + n) % n;
  int i1 = i0 + 1;
  if (i1 >= n) i1 -= n;
  c.i0 = i0;
  c.i1 = i1;
  return c;
}

// adj_a(c) += sum_p W(x_p - c) val_a(p): the transpose of the readout with
// respect to the mesh (global FP32 atomics; meshes cleared by the caller).
__global__ void k_paint_values3(const float* __restrict__ x, const float* __restrict__ val,
                                int64_t stride, int64_t np, float* __restrict__ a0,
                                float* __restrict__ a1, float* __restrict__ a2, int n, int ncz,
                                float dx_inv) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= np) return;
  const Cic1 cx = cic1(x[i], dx_inv, n);
  const Cic1 cy = cic1(x[i + stride], dx_inv, n);
  const Cic1 cz = cic1(x[i + 2 * stride], dx_inv, n);
  const float v[3] = {val[i], val[i + stride], val[i + 2 * stride]};
  float* out[3] = {a0, a1, a2};
#pragma unroll
  for (int ex = 0; ex < 2; ++ex)
#pragma unroll
    for (int ey = 0; ey < 2; ++ey)
#pragma unroll
      for (int ez = 0; ez < 2; ++ez) {
        const float w = (ex ? cx.w1 : cx.w0) * (ey ? cy.w1 : cy.w0) * (ez ? cz.w1 : cz.w0);
        const int64_t c = ((int64_t)(ex ? cx.i1 : cx.i0) * n + (ey ? cy.i1 : cy.i0)) * ncz +
                          (ez ? cz.i1 : cz.i0);
#pragma unroll
        for (int a = 0; a < 3; ++a) unsafeAtomicAdd(out[a] + c, w * v[a]);
      }
}

// xbar_p += sum_c grad W(x_p - c) S_p(c),  S_p(c) = fac rhobar(c)
//           + sum_a fbar_a(p) F_a(c)   (F_a terms optional)
__global__ void k_grad_readout(const float* __restrict__ x, int64_t stride, int64_t np,
                               const float* __restrict__ rhobar, float fac,
                               const float* __restrict__ f0, const float* __restrict__ f1,
                               const float* __restrict__ f2, const float* __restrict__ fbar,
                               float* __restrict__ xbar, int n, int ncz, float dx_inv) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= np) return;
  const Cic1 cx = cic1(x[i], dx_inv, n);
  const Cic1 cy = cic1(x[i + stride], dx_inv, n);
  const Cic1 cz = cic1(x[i + 2 * stride], dx_inv, n);
  float fb[3] = {0.f, 0.f, 0.f};
  const bool forces = fbar != nullptr;
  if (forces) {
    fb[0] = fbar[i];
    fb[1] = fbar[i + stride];
    fb[2] = fbar[i + 2 * stride];
  }
  float gx = 0.f, gy = 0.f, gz = 0.f;
#pragma unroll
  for (int ex = 0; ex < 2; ++ex)
#pragma unroll
    for (int ey = 0; ey < 2; ++ey)
#pragma unroll
      for (int ez = 0; ez < 2; ++ez) {
        const int64_t c = ((int64_t)(ex ? cx.i1 : cx.i0) * n + (ey ? cy.i1 : cy.i0)) * ncz +
                          (ez ? cz.i1 : cz.i0);
        float S = fac * rhobar[c];
        if (forces) S += fb[0] * f0[c] + fb[1] * f1[c] + fb[2] * f2[c];
        const float wx = ex ? cx.w1 : cx.w0, wy = ey ? cy.w1 : cy.w0, wz = ez ? cz.w1 : cz.w0;
        const float sx = ex ? dx_inv : -dx_inv, sy = ey ? dx_inv : -dx_inv,
                    sz = ez ? dx_inv : -dx_inv;
        gx += sx * wy * wz * S;
        gy += wx * sy * wz * S;
        gz += wx * wy * sz * S;
      }
  xbar[i] += gx;
  xbar[i + stride] += gy;
  xbar[i + 2 * stride] += gz;
}

// One COLA step backwards (py/cola.cpp:79-88, 117-124 transposed):
//   vtot = vbar + dt xbar; dx1bar += da1 xbar + c q1 vtot;
//   dx2bar += da2 xbar + c q2 vtot; fbar = c vtot; vbar = vtot
__global__ void k_cola_step_adjoint(const float* __restrict__ xbar, float* __restrict__ vbar,
                                    float* __restrict__ dx1bar, float* __restrict__ dx2bar,
                                    float* __restrict__ fbar, int64_t n3, float dt, float da1,
                                    float da2, float c, float q1, float q2) {
  const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n3) return;
  const float xb = xbar[j];
  const float vt = vbar[j] + dt * xb;
  dx1bar[j] += da1 * xb + c * q1 * vt;
  dx2bar[j] += da2 * xb + c * q2 * vt;
  fbar[j] = c * vt;
  vbar[j] = vt;
}

__global__ void k_axpy(float* __restrict__ y, const float* __restrict__ x, float a, int64_t n) {
  const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j < n) y[j] += a * x[j];
}

// particle plane (np = n^3, lattice order) <-> padded grid
__global__ void k_plane_to_grid(const float* __restrict__ plane, float* __restrict__ grid, int n,
                                int ncz) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (int64_t)n * n * n) return;
  const int64_t row = i / n;
  grid[row * ncz + (i - row * n)] = plane[i];
}

// psibar_ij from divbar (lpt.cpp:527-532 product rule), in place over psi:
// d/dP11 = P22 + P33, d/dP22 = P11 + P33, d/dP33 = P11 + P22,
// d/dP12 = -2 P12, d/dP13 = -2 P13, d/dP23 = -2 P23.
__global__ void k_lpt_div2_adjoint(const float* __restrict__ divbar, float* p0, float* p1,
                                   float* p2, float* p3, float* p4, float* p5, int n, int ncz) {
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= (int64_t)n * n * n) return;
  const int64_t row = c / n;
  const int64_t idx = row * ncz + (c - row * n);
  const float d = divbar[idx];
  const float a0 = p0[idx], a1 = p1[idx], a2 = p2[idx], a3 = p3[idx], a4 = p4[idx], a5 = p5[idx];
  p0[idx] = d * (a3 + a5);
  p1[idx] = -2.f * d * a1;
  p2[idx] = -2.f * d * a2;
  p3[idx] = d * (a0 + a5);
  p4[idx] = -2.f * d * a4;
  p5[idx] = d * (a0 + a3);
}

struct LptMode {
  double k[3];
  double k2;
  int k2i;
  bool nyq;
  int64_t idx;
};
__device__ __forceinline__ LptMode lpt_mode(int64_t mi, int n, int cpitch, double kf) {
  LptMode m;
  const int nh = n / 2 + 1;
  const int iz = (int)(mi % nh);
  const int64_t rxy = mi / nh;
  const int iy = (int)(rxy % n), ix = (int)(rxy / n);
  const int h = n / 2;
  const int kx = ix < h ? ix : ix - n, ky = iy < h ? iy : iy - n, kz = iz < h ? iz : iz - n;
  m.k[0] = kf * kx;
  m.k[1] = kf * ky;
  m.k[2] = kf * kz;
  m.k2 = m.k[0] * m.k[0] + m.k[1] * m.k[1] + m.k[2] * m.k[2];
  m.k2i = kx * kx + ky * ky + kz * kz;
  m.nyq = ix == h || iy == h || iz == h;
  m.idx = rxy * cpitch + iz;
  return m;
}

// divbar_k = sum_a (i k_a / k^2) G_a(k) / n^3  (transpose of the Psi2 map,
// lpt.cpp:579-583), written into g0's spectrum.
__global__ void k_lpt_divbar_k(float2* g0, const float2* __restrict__ g1,
                               const float2* __restrict__ g2, int n, int cpitch, double kf,
                               double inv_n3) {
  const int64_t total = (int64_t)n * n * (n / 2 + 1);
  const int64_t mi = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (mi >= total) return;
  const LptMode m = lpt_mode(mi, n, cpitch, kf);
  if (m.k2i == 0) {
    g0[m.idx] = make_float2(0.f, 0.f);
    return;
  }
  const float2 v[3] = {g0[m.idx], g1[m.idx], g2[m.idx]};
  double re = 0.0, im = 0.0;
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const double c = m.k[a] / m.k2 * inv_n3;
    re += -c * v[a].y;
    im += c * v[a].x;
  }
  g0[m.idx] = make_float2((float)re, (float)im);
}

// wbar_k = amp(|k|) mask(k) [ sum_a (-i k_a/k^2) H_a(k) + sum_ij (-k_i k_j/k^2) P_ij(k) ]
// (transposes of w -> dx1 and w -> psi_ij), written into h0's spectrum.
__global__ void k_lpt_noisebar_k(float2* h0, const float2* __restrict__ h1,
                                 const float2* __restrict__ h2, const float2* __restrict__ q0,
                                 const float2* __restrict__ q1, const float2* __restrict__ q2,
                                 const float2* __restrict__ q3, const float2* __restrict__ q4,
                                 const float2* __restrict__ q5, int n, int cpitch, double kf,
                                 const double* __restrict__ amp) {
  const int64_t total = (int64_t)n * n * (n / 2 + 1);
  const int64_t mi = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (mi >= total) return;
  const LptMode m = lpt_mode(mi, n, cpitch, kf);
  if (m.nyq || m.k2i == 0) {
    h0[m.idx] = make_float2(0.f, 0.f);
    return;
  }
  const double A = amp[m.k2i];
  const float2 h[3] = {h0[m.idx], h1[m.idx], h2[m.idx]};
  const float2 q[6] = {q0[m.idx], q1[m.idx], q2[m.idx], q3[m.idx], q4[m.idx], q5[m.idx]};
  const int pi[6] = {0, 0, 0, 1, 1, 2};
  const int pj[6] = {0, 1, 2, 1, 2, 2};
  double re = 0.0, im = 0.0;
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const double c = -m.k[a] / m.k2;   // -i c' with c' = k_a/k^2: (x + iy)(-i c') = (c' y, -c' x)
    re += -c * h[a].y;
    im += c * h[a].x;
  }
#pragma unroll
  for (int s = 0; s < 6; ++s) {
    const double c = -m.k[pi[s]] * m.k[pj[s]] / m.k2;
    re += c * q[s].x;
    im += c * q[s].y;
  }
  h0[m.idx] = make_float2((float)(A * re), (float)(A * im));
}

}  // namespace

void adj_paint_values3(const float* x, const float* val, int64_t stride, int64_t np,
                       float* const out[3], int n, float boxsize) {
  if (np == 0) return;
  launch("adj_paint_values3", (double)np * 24.0, k_paint_values3, dim3(blocks_for(np)),
         dim3(kBlock), 0, x, val, stride, np, out[0], out[1], out[2], n, grid_rpitch(n),
         (float)n / boxsize);
}

void adj_grad_readout(const float* x, int64_t stride, int64_t np, const float* rhobar, float fac,
                      const float* const F[3], const float* fbar, float* xbar, int n,
                      float boxsize) {
  if (np == 0) return;
  launch("adj_grad_readout", (double)np * 48.0 + 16.0 * (double)n * n * n, k_grad_readout,
         dim3(blocks_for(np)), dim3(kBlock), 0, x, stride, np, rhobar, fac,
         F ? F[0] : nullptr, F ? F[1] : nullptr, F ? F[2] : nullptr, fbar, xbar, n,
         grid_rpitch(n), (float)n / boxsize);
}

void adj_cola_step(const float* xbar, float* vbar, float* dx1bar, float* dx2bar, float* fbar,
                   int64_t n3, float dt, float da1, float da2, float c, float q1, float q2) {
  launch("adj_cola_step", (double)n3 * 32.0, k_cola_step_adjoint, dim3(blocks_for(n3)),
         dim3(kBlock), 0, xbar, vbar, dx1bar, dx2bar, fbar, n3, dt, da1, da2, c, q1, q2);
}

void adj_axpy(float* y, const float* x, float a, int64_t n) {
  launch("adj_axpy", (double)n * 12.0, k_axpy, dim3(blocks_for(n)), dim3(kBlock), 0, y, x, a, n);
}

void plane_to_grid(const float* plane, float* grid, int n) {
  const int64_t n3 = (int64_t)n * n * n;
  launch("plane_to_grid", (double)n3 * 8.0, k_plane_to_grid, dim3(blocks_for(n3)), dim3(kBlock),
         0, plane, grid, n, grid_rpitch(n));
}

void adj_lpt_div2(const float* divbar, float* const psi[6], int n) {
  const int64_t n3 = (int64_t)n * n * n;
  launch("adj_lpt_div2", (double)n3 * 52.0, k_lpt_div2_adjoint, dim3(blocks_for(n3)),
         dim3(kBlock), 0, divbar, psi[0], psi[1], psi[2], psi[3], psi[4], psi[5], n,
         grid_rpitch(n));
}

void adj_lpt_divbar_k(float* const g[3], int n, double boxsize) {
  const int64_t total = (int64_t)n * n * (n / 2 + 1);
  launch("adj_lpt_divbar_k", (double)total * 32.0, k_lpt_divbar_k, dim3(blocks_for(total)),
         dim3(kBlock), 0, (float2*)g[0], (const float2*)g[1], (const float2*)g[2], n,
         grid_cpitch(n), 2.0 * M_PI / boxsize, 1.0 / ((double)n * n * n));
}

void adj_lpt_noisebar_k(float* const h[3], float* const q[6], int n, double boxsize,
                        const double* amp) {
  const int64_t total = (int64_t)n * n * (n / 2 + 1);
  launch("adj_lpt_noisebar_k", (double)total * 80.0, k_lpt_noisebar_k, dim3(blocks_for(total)),
         dim3(kBlock), 0, (float2*)h[0], (const float2*)h[1], (const float2*)h[2],
         (const float2*)q[0], (const float2*)q[1], (const float2*)q[2], (const float2*)q[3],
         (const float2*)q[4], (const float2*)q[5], n, grid_cpitch(n), 2.0 * M_PI / boxsize, amp);
}

}  // namespace fs2
