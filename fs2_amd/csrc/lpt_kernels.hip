// 2LPT initial-condition kernels for gfx950 (py/lpt.cpp:100-225, 439-591).
//
// The Gaussian field is white noise from Philox4x32-10 (Salmon et al.,
// SC'11) + Box-Muller, keyed by the seed and counted by cell, replacing the
// N-GenIC / GSL ranlxd1 draw of py/lpt.cpp:235-329.  Mode arithmetic is FP64
// as in the reference; fields are stored FP32.
#include <cmath>

#include "fs2_internal.h"

#pragma clang fp contract(off)

namespace fs2 {
namespace {

constexpr int kBlock = 256;
inline unsigned blocks_for(int64_t n) { return (unsigned)((n + kBlock - 1) / kBlock); }

__device__ __forceinline__ void philox4x32_10(uint32_t c[4], uint32_t k0, uint32_t k1) {
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    const uint32_t lo0 = 0xD2511F53u * c[0], hi0 = __umulhi(0xD2511F53u, c[0]);
    const uint32_t lo1 = 0xCD9E8D57u * c[2], hi1 = __umulhi(0xCD9E8D57u, c[2]);
    const uint32_t n0 = hi1 ^ c[1] ^ k0, n2 = hi0 ^ c[3] ^ k1;
    c[0] = n0; c[1] = lo1; c[2] = n2; c[3] = lo0;
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
}

// Call c fills cells 4c..4c+3 (linear index (ix*n + iy)*n + iz); n is even,
// so n^3 is a multiple of 4 and every cell gets exactly one value.
__global__ void k_white_noise(float* __restrict__ g, int n, int ncz, int64_t ncall, uint32_t s0,
                              uint32_t s1) {
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ncall) return;
  uint32_t r[4] = {(uint32_t)c, (uint32_t)((uint64_t)c >> 32), 0u, 0u};
  philox4x32_10(r, s0, s1);
  const double sc = 2.3283064365386962890625e-10;   // 2^-32
  const double u0 = ((double)r[0] + 0.5) * sc, u1 = ((double)r[1] + 0.5) * sc;
  const double u2 = ((double)r[2] + 0.5) * sc, u3 = ((double)r[3] + 0.5) * sc;
  const double ra = sqrt(-2.0 * log(u0)), rb = sqrt(-2.0 * log(u2));
  double sa, ca, sb, cb;
  sincospi(2.0 * u1, &sa, &ca);
  sincospi(2.0 * u3, &sb, &cb);
  const float v[4] = {(float)(ra * ca), (float)(ra * sa), (float)(rb * cb), (float)(rb * sb)};
  // the four values are cells 4c..4c+3; with n % 4 == 2 they may span two rows
  const int64_t i = 4 * c;
  int64_t row = i / n;
  int iz = (int)(i - row * n);
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    g[row * ncz + iz] = v[j];
    if (++iz == n) {
      iz = 0;
      ++row;
    }
  }
}

struct Mode {
  int ix, iy, iz;
  double k[3];
  double k2;
  int k2i;
  bool nyq;
};

// Wave vector of half-complex index idx, lpt.cpp convention: an index below
// n/2 is positive, otherwise index - n (so n/2 maps to -n/2, lpt.cpp:469-482).
__device__ __forceinline__ Mode mode_of(int64_t idx, int n, double dk) {
  Mode m;
  const int nh = n / 2 + 1;
  m.iz = (int)(idx % nh);
  const int64_t rxy = idx / nh;
  m.iy = (int)(rxy % n);
  m.ix = (int)(rxy / n);
  const int h = n / 2;
  const int kx = m.ix < h ? m.ix : m.ix - n;
  const int ky = m.iy < h ? m.iy : m.iy - n;
  const int kz = m.iz < h ? m.iz : m.iz - n;
  m.k[0] = dk * kx;
  m.k[1] = dk * ky;
  m.k[2] = dk * kz;
  m.k2 = m.k[0] * m.k[0] + m.k[1] * m.k[1] + m.k[2] * m.k[2];
  m.k2i = kx * kx + ky * ky + kz * kz;
  m.nyq = (m.ix == h) || (m.iy == h) || (m.iz == h);
  return m;
}

// delta_k = W_k amp(|k|) with Nyquist planes and k = 0 zeroed (lpt.cpp:328-331);
// Psi1_i = i k_i/k^2 delta_k (lpt.cpp:374-375);
// Psi_ij = i k_j Psi1_i for ij = 11,12,13,22,23,33 (lpt.cpp:486-502).
__global__ void k_psi_ij(float2* __restrict__ dk, float2* __restrict__ o0, float2* __restrict__ o1,
                         float2* __restrict__ o2, float2* __restrict__ o3, float2* __restrict__ o4,
                         float2* __restrict__ o5, int n, int cpitch, double kf,
                         const double* __restrict__ amp) {
  const int64_t total = (int64_t)n * n * (n / 2 + 1);
  const int64_t mi = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (mi >= total) return;
  const Mode m = mode_of(mi, n, kf);
  const int64_t idx = ((int64_t)m.ix * n + m.iy) * cpitch + m.iz;
  float2 d = make_float2(0.f, 0.f);
  if (!m.nyq && m.k2i > 0) {
    const float2 w = dk[idx];
    const double a = amp[m.k2i];
    d = make_float2((float)(w.x * a), (float)(w.y * a));
  }
  dk[idx] = d;
  float2 psi[3];
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    const double c = m.k2 > 0 ? m.k[i] / m.k2 : 0.0;
    psi[i] = make_float2((float)(-c * d.y), (float)(c * d.x));
  }
  float2* outs[6] = {o0, o1, o2, o3, o4, o5};
  const int pi[6] = {0, 0, 0, 1, 1, 2};
  const int pj[6] = {0, 1, 2, 1, 2, 2};
#pragma unroll
  for (int q = 0; q < 6; ++q) {
    const float2 p = psi[pi[q]];
    const double kj = m.k[pj[q]];
    outs[q][idx] = make_float2((float)(-(double)p.y * kj), (float)((double)p.x * kj));
  }
}

// div Psi2 = P11 (P22 + P33) + P22 P33 - P12^2 - P13^2 - P23^2 (lpt.cpp:527-532)
__global__ void k_div2(float* out, const float* p0, const float* p1, const float* p2,
                       const float* p3, const float* p4, const float* p5, int n, int ncz) {
  const int64_t cells = (int64_t)n * n * n;
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= cells) return;
  const int64_t row = c / n;
  const int64_t idx = row * ncz + (c - row * n);
  const float a0 = p0[idx], a1 = p1[idx], a2 = p2[idx], a3 = p3[idx], a4 = p4[idx], a5 = p5[idx];
  out[idx] = a0 * (a3 + a5) + a3 * a5 - a1 * a1 - a2 * a2 - a4 * a4;
}

// Psi1_k from delta_k and Psi2_k = -i k/k^2 D_k (lpt.cpp:579-583), k = 0 -> 0.
// delta_k (in/out) receives Psi1_z; out: Psi1_x, Psi1_y, Psi2_x, Psi2_y, Psi2_z.
__global__ void k_psi12(float2* __restrict__ delta, const float2* __restrict__ divk,
                        float2* __restrict__ q0, float2* __restrict__ q1, float2* __restrict__ q2,
                        float2* __restrict__ q3, float2* __restrict__ q4, int n, int cpitch,
                        double kf) {
  const int64_t total = (int64_t)n * n * (n / 2 + 1);
  const int64_t mi = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (mi >= total) return;
  const Mode m = mode_of(mi, n, kf);
  const int64_t idx = ((int64_t)m.ix * n + m.iy) * cpitch + m.iz;
  const float2 d = delta[idx];
  const float2 D = divk[idx];
  float2 p1[3], p2[3];
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    const double c = m.k2 > 0 ? m.k[i] / m.k2 : 0.0;
    p1[i] = make_float2((float)(-c * d.y), (float)(c * d.x));
    p2[i] = make_float2((float)((double)D.y * m.k[i] / (m.k2 > 0 ? m.k2 : 1.0)),
                        (float)(-(double)D.x * m.k[i] / (m.k2 > 0 ? m.k2 : 1.0)));
    if (m.k2i == 0) p2[i] = make_float2(0.f, 0.f);
  }
  q0[idx] = p1[0];
  q1[idx] = p1[1];
  delta[idx] = p1[2];
  q2[idx] = p2[0];
  q3[idx] = p2[1];
  q4[idx] = p2[2];
}

// Particle grid and displacements, lpt.cpp:182-211.
__global__ void k_lattice(LatticeArgs o, const float* __restrict__ s0, const float* __restrict__ s1,
                          const float* __restrict__ s2, const float* __restrict__ t0,
                          const float* __restrict__ t1, const float* __restrict__ t2, int n,
                          int ncz, float dx, float offset, float D1, float D2, float Dv, float D2v,
                          double nmesh3_inv) {
  const int64_t np = (int64_t)n * n * n;
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= np) return;
  const int iz = (int)(i % n);
  const int64_t rxy = i / n;
  const int iy = (int)(rxy % n);
  const int ix = (int)(rxy / n);
  const int64_t g = rxy * ncz + iz;
  const float q[3] = {((float)ix + offset) * dx, ((float)iy + offset) * dx,
                      ((float)iz + offset) * dx};
  const float* ps1[3] = {s0, s1, s2};
  const float* ps2[3] = {t0, t1, t2};
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const float dis = ps1[k][g];
    const float dis2 = (float)(nmesh3_inv * (double)ps2[k][g]);
    o.x[i + k * o.stride] = q[k] + D1 * dis + D2 * dis2;
    o.v[i + k * o.stride] = Dv * dis + D2v * dis2;
    o.dx1[i + k * o.stride] = dis;
    o.dx2[i + k * o.stride] = dis2;
  }
  if (o.id) o.id[i] = (uint64_t)i + 1;
}

}  // namespace

void white_noise(float* grid, int n, uint64_t seed) {
  if (n % 2 != 0) fail(FS2_ERR_VALUE, "white noise needs an even nc (nc = %d)", n);
  const int64_t ncall = (int64_t)n * n * n / 4;
  launch("white_noise", (double)n * n * n * 4.0, k_white_noise, dim3(blocks_for(ncall)),
         dim3(kBlock), 0, grid, n, grid_rpitch(n), ncall, (uint32_t)(seed & 0xffffffffu),
         (uint32_t)(seed >> 32));
}

void lpt_psi_ij(float* dk, float* const out[6], int n, double boxsize, const double* amp) {
  const int64_t total = (int64_t)n * n * (n / 2 + 1);
  launch("lpt_psi_ij", (double)total * 8.0 * 8.0, k_psi_ij, dim3(blocks_for(total)), dim3(kBlock),
         0, (float2*)dk, (float2*)out[0], (float2*)out[1], (float2*)out[2], (float2*)out[3],
         (float2*)out[4], (float2*)out[5], n, grid_cpitch(n), 2.0 * M_PI / boxsize, amp);
}

void lpt_div2(float* out, const float* const psi[6], int n) {
  const int64_t cells = (int64_t)n * n * n;
  launch("lpt_div2", (double)cells * 28.0, k_div2, dim3(blocks_for(cells)), dim3(kBlock), 0, out,
         psi[0], psi[1], psi[2], psi[3], psi[4], psi[5], n, grid_rpitch(n));
}

void lpt_psi12(float* psi1_z_inout, const float* div_k, float* const out5[5], int n,
               double boxsize, const double* amp) {
  (void)amp;
  const int64_t total = (int64_t)n * n * (n / 2 + 1);
  launch("lpt_psi12", (double)total * 8.0 * 8.0, k_psi12, dim3(blocks_for(total)), dim3(kBlock), 0,
         (float2*)psi1_z_inout, (const float2*)div_k, (float2*)out5[0], (float2*)out5[1],
         (float2*)out5[2], (float2*)out5[3], (float2*)out5[4], n, grid_cpitch(n),
         2.0 * M_PI / boxsize);
}

void lpt_lattice(LatticeArgs out, const float* const psi1[3], const float* const psi2[3], int n,
                 float dx, float offset, float D1, float D2, float Dv, float D2v,
                 double nmesh3_inv) {
  const int64_t np = (int64_t)n * n * n;
  launch("lpt_lattice", (double)np * (24.0 + 56.0), k_lattice, dim3(blocks_for(np)), dim3(kBlock),
         0, out, psi1[0], psi1[1], psi1[2], psi2[0], psi2[1], psi2[2], n, grid_rpitch(n), dx,
         offset, D1, D2, Dv, D2v, nmesh3_inv);
}

}  // namespace fs2
