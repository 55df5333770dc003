// Particle <-> mesh kernels of the PM solver for gfx950.
//
// CIC arithmetic is kept in the reference's FP32 operation order
// (py/pm.cpp:84-133, 155-218), so floating-point contraction is disabled in
// this file: x*dx_inv, the (int) cast, w1 = x - i0, w0 = 1 - w1 and
// fac*wx*wy*wz are each rounded exactly as the C++ reference rounds them.
#include <algorithm>
#include <cmath>
#include <cstdio>

#include "fs2_internal.h"

#pragma clang fp contract(off)

// Bin sums of the allocation-free P(k) entry point
// (fs2_pm_power_spectrum_noalloc): counts, k sums and P sums of up to
// FS2_PK_NOALLOC_MAX_BINS bins, FP64.  Each GPU holds its own copy for the
// life of the process.  In host code this name is only a stand-in for the
// device address: host code uses the pointer that pk_scratch_init obtains
// with hipGetSymbolAddress.
static __device__ __align__(256) double scratch[3 * FS2_PK_NOALLOC_MAX_BINS];

namespace fs2 {
namespace {

constexpr int kBlock = 256;
constexpr int kMaxDevices = 64;
double* g_pk_scratch[kMaxDevices];   // device address of `scratch`, per device

inline unsigned blocks_for(int64_t n, int per_block = kBlock) {
  return (unsigned)((n + per_block - 1) / per_block);
}

// py/util.h:14-24 periodic_wrapup_p
__device__ __forceinline__ float wrap1(float x, float L) {
  if (x < 0.f) x += L;
  if (x >= L) x -= L;
  return x;
}

// Lower cell index and weights of one coordinate, py/pm.cpp:85-115.
// An index outside [0, n) (only reachable from positions the reference would
// also mishandle: not wrapped, NaN) is folded back into the mesh so that no
// access leaves the grid.
struct Cic1 {
  int i0, i1;
  float w0, w1;
};
__device__ __forceinline__ Cic1 cic1(float x, float dx_inv, int n) {
  Cic1 c;
  const float X = x * dx_inv;
  int i0 = (int)X;                     // truncation, pm.cpp:95
  c.w1 = X - (float)i0;                // pm.cpp:100
  c.w0 = 1.f - c.w1;                   // pm.cpp:105
  if (i0 >= n) i0 = 0;                 // pm.cpp:109
  if ((unsigned)i0 >= (unsigned)n) i0 = ((i0 % n) + n) % n;
  int i1 = i0 + 1;
  if (i1 >= n) i1 -= n;                // pm.cpp:113
  c.i0 = i0;
  c.i1 = i1;
  return c;
}

__global__ void k_fill(float4* __restrict__ g, int64_t n4, float v) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n4) g[i] = make_float4(v, v, v, v);
}

__global__ void k_wrap(float* __restrict__ x, int64_t stride, int64_t np, float L) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= np) return;
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const float v = x[i + a * stride];
    const float w = wrap1(v, L);
    if (w != v) x[i + a * stride] = w;
  }
}

// pm_assign_cic_density (py/pm.cpp:57-139) with the periodic wrap of
// pm_domain_send_positions (py/pm_domain.cpp:115) fused into the load.
__global__ void k_paint_atomic(float* __restrict__ x, int64_t stride, int64_t np,
                               float* __restrict__ grid, int n, int ncz, float dx_inv,
                               float fac, float L, int do_wrap) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= np) return;
  float p[3];
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    p[a] = x[i + a * stride];
    if (do_wrap) {
      const float w = wrap1(p[a], L);
      if (w != p[a]) x[i + a * stride] = w;
      p[a] = w;
    }
  }
  const Cic1 cx = cic1(p[0], dx_inv, n);
  const Cic1 cy = cic1(p[1], dx_inv, n);
  const Cic1 cz = cic1(p[2], dx_inv, n);
  const int64_t r00 = ((int64_t)cx.i0 * n + cy.i0) * ncz;
  const int64_t r01 = ((int64_t)cx.i0 * n + cy.i1) * ncz;
  const int64_t r10 = ((int64_t)cx.i1 * n + cy.i0) * ncz;
  const int64_t r11 = ((int64_t)cx.i1 * n + cy.i1) * ncz;
  // fac*wx*wy*wz, left to right (pm.cpp:122-132)
  unsafeAtomicAdd(grid + r00 + cz.i0, fac * cx.w0 * cy.w0 * cz.w0);
  unsafeAtomicAdd(grid + r00 + cz.i1, fac * cx.w0 * cy.w0 * cz.w1);
  unsafeAtomicAdd(grid + r01 + cz.i0, fac * cx.w0 * cy.w1 * cz.w0);
  unsafeAtomicAdd(grid + r01 + cz.i1, fac * cx.w0 * cy.w1 * cz.w1);
  unsafeAtomicAdd(grid + r10 + cz.i0, fac * cx.w1 * cy.w0 * cz.w0);
  unsafeAtomicAdd(grid + r10 + cz.i1, fac * cx.w1 * cy.w0 * cz.w1);
  unsafeAtomicAdd(grid + r11 + cz.i0, fac * cx.w1 * cy.w1 * cz.w0);
  unsafeAtomicAdd(grid + r11 + cz.i1, fac * cx.w1 * cy.w1 * cz.w1);
}

// force_at_particle_locations (py/pm.cpp:142-220) for all three axes in one
// pass: the CIC stencil is computed once and 24 values gathered.
__global__ void k_readout3(const float* __restrict__ x, int64_t stride, int64_t np,
                           const float* __restrict__ g0, const float* __restrict__ g1,
                           const float* __restrict__ g2, float* __restrict__ f, int n,
                           int ncz, float dx_inv) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= np) return;
  const Cic1 cx = cic1(x[i], dx_inv, n);
  const Cic1 cy = cic1(x[i + stride], dx_inv, n);
  const Cic1 cz = cic1(x[i + 2 * stride], dx_inv, n);
  const int64_t r00 = ((int64_t)cx.i0 * n + cy.i0) * ncz;
  const int64_t r01 = ((int64_t)cx.i0 * n + cy.i1) * ncz;
  const int64_t r10 = ((int64_t)cx.i1 * n + cy.i0) * ncz;
  const int64_t r11 = ((int64_t)cx.i1 * n + cy.i1) * ncz;
  const float* gs[3] = {g0, g1, g2};
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const float* g = gs[a];
    // pm.cpp:203-218: f = 0; f += (ix0 terms); f += (ix1 terms)
    float s = 0.f;
    s += g[r00 + cz.i0] * cx.w0 * cy.w0 * cz.w0 + g[r00 + cz.i1] * cx.w0 * cy.w0 * cz.w1 +
         g[r01 + cz.i0] * cx.w0 * cy.w1 * cz.w0 + g[r01 + cz.i1] * cx.w0 * cy.w1 * cz.w1;
    s += g[r10 + cz.i0] * cx.w1 * cy.w0 * cz.w0 + g[r10 + cz.i1] * cx.w1 * cy.w0 * cz.w1 +
         g[r11 + cz.i0] * cx.w1 * cy.w1 * cz.w0 + g[r11 + cz.i1] * cx.w1 * cy.w1 * cz.w1;
    f[i + a * stride] = s;
  }
}

// FP64 sum over the n^3 cells of a padded grid (py/pm.cpp:364-370).
__global__ void k_grid_sum(const float* __restrict__ g, int n, int ncz, double* __restrict__ out) {
  __shared__ double red[kBlock];
  const int64_t rows = (int64_t)n * n;
  double acc = 0.0;
  for (int64_t r = blockIdx.x; r < rows; r += gridDim.x) {
    const float* row = g + r * ncz;
    for (int z = threadIdx.x; z < n; z += blockDim.x) acc += (double)row[z];
  }
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) atomicAdd(out, red[0]);
}

// Binned P(k) over the half-complex modes; LDS bins, FP64 sums.
__global__ void k_power_spectrum(const float2* __restrict__ dk, int n, int cpitch, int nbins,
                                 double dk_bin, double pnorm, double* __restrict__ out /* 3*nbins */) {
  extern __shared__ double bins[];   // 3*nbins
  for (int b = threadIdx.x; b < 3 * nbins; b += blockDim.x) bins[b] = 0.0;
  __syncthreads();
  const int nh = n / 2 + 1;
  const int64_t total = (int64_t)n * n * nh;
  for (int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < total;
       idx += (int64_t)gridDim.x * blockDim.x) {
    const int iz = (int)(idx % nh);
    const int64_t rxy = idx / nh;
    const int iy = (int)(rxy % n);
    const int ix = (int)(rxy / n);
    const int kx = ix <= n / 2 ? ix : ix - n;
    const int ky = iy <= n / 2 ? iy : iy - n;
    const double kk = sqrt((double)kx * kx + (double)ky * ky + (double)iz * iz);
    if (kk == 0.0) continue;
    const int b = (int)floor(kk / dk_bin);
    if (b >= nbins) continue;
    const double w = (iz == 0 || iz == n / 2) ? 1.0 : 2.0;
    const float2 v = dk[rxy * cpitch + iz];
    const double p = ((double)v.x * v.x + (double)v.y * v.y) * pnorm;
    atomicAdd(&bins[b], w);
    atomicAdd(&bins[nbins + b], w * kk);
    atomicAdd(&bins[2 * nbins + b], w * p);
  }
  __syncthreads();
  for (int b = threadIdx.x; b < 3 * nbins; b += blockDim.x)
    if (bins[b] != 0.0) atomicAdd(&out[b], bins[b]);
}

__global__ void k_zero_f64(double* __restrict__ p, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0.0;
}

// Clears the 3*nbins sums with a kernel (no memset node, so the sequence can
// be captured), bins them, copies counts, k sums and P sums into the caller's
// arrays and normalises those in place: no host or device allocation.
void power_spectrum_sums(const float* delta_k, int n, double boxsize, int nbins, double dk_bin,
                         double* d, double* k, double* P, double* nmodes) {
  launch("pk_clear", (double)nbins * 24.0, k_zero_f64, dim3((3 * nbins + kBlock - 1) / kBlock),
         dim3(kBlock), 0, d, 3 * nbins);
  const double ng = (double)n * n * n;
  const double pnorm = boxsize * boxsize * boxsize / (ng * ng);
  const int64_t total = (int64_t)n * n * (n / 2 + 1);
  const unsigned nb = (unsigned)std::min<int64_t>(blocks_for(total), 2048);
  launch("power_spectrum", (double)total * 8.0, k_power_spectrum, dim3(nb), dim3(kBlock),
         3 * nbins * sizeof(double), (const float2*)delta_k, n, grid_cpitch(n), nbins, dk_bin,
         pnorm, d);
  const size_t bytes = sizeof(double) * nbins;
  FS2_HIP(hipMemcpyAsync(nmodes, d, bytes, hipMemcpyDeviceToHost, stream()));
  FS2_HIP(hipMemcpyAsync(k, d + nbins, bytes, hipMemcpyDeviceToHost, stream()));
  FS2_HIP(hipMemcpyAsync(P, d + 2 * nbins, bytes, hipMemcpyDeviceToHost, stream()));
  FS2_HIP(hipStreamSynchronize(stream()));
  const double kf = 2.0 * M_PI / boxsize;
  for (int b = 0; b < nbins; ++b) {
    const double c = nmodes[b];
    k[b] = c > 0 ? k[b] / c * kf : 0.0;
    P[b] = c > 0 ? P[b] / c : 0.0;
  }
}

}  // namespace

void pk_scratch_init() {
  int dev = 0;
  FS2_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) fail(FS2_ERR_DEVICE, "device %d out of range", dev);
  void* ptr = nullptr;
  const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
  if (e != hipSuccess || ptr == nullptr) {
    (void)hipGetLastError();
    fail(FS2_ERR_DEVICE, "hipGetSymbolAddress(scratch) failed: %s", hipGetErrorString(e));
  }
  g_pk_scratch[dev] = (double*)ptr;
}

void fill_grid(float* grid, int n, float value) {
  const int64_t total = (int64_t)grid_floats(n);
  const int64_t n4 = total / 4;
  launch("fill_grid", (double)total * 4.0, k_fill, dim3(blocks_for(n4)), dim3(kBlock), 0,
         (float4*)grid, n4, value);
}

void wrap_positions(PosArgs p, float boxsize) {
  if (p.np == 0) return;
  launch("wrap", (double)p.np * 24.0, k_wrap, dim3(blocks_for(p.np)), dim3(kBlock), 0, p.x,
         p.stride, p.np, boxsize);
}

void paint_atomic(PosArgs p, float* grid, int n, float boxsize, float fac, bool wrap) {
  if (p.np == 0) return;
  const int ncz = grid_rpitch(n);
  // pm.cpp:77: Float dx_inv = nc/boxsize
  const float dx_inv = (float)n / boxsize;
  // algorithmic bytes: read x (12 B/particle); the mesh write is counted
  // with the clear that precedes it
  launch("paint_atomic", (double)p.np * 12.0, k_paint_atomic, dim3(blocks_for(p.np)),
         dim3(kBlock), 0, p.x, p.stride, p.np, grid, n, ncz, dx_inv, fac, boxsize,
         wrap ? 1 : 0);
}

void readout3(PosArgs p, const float* g0, const float* g1, const float* g2, float* f, int n,
              float boxsize) {
  if (p.np == 0) return;
  const int ncz = grid_rpitch(n);
  const float dx_inv = (float)n / boxsize;
  const double bytes = (double)p.np * 24.0 + 3.0 * 4.0 * (double)n * n * n;
  launch("readout3", bytes, k_readout3, dim3(blocks_for(p.np)), dim3(kBlock), 0,
         (const float*)p.x, p.stride, p.np, g0, g1, g2, f, n, ncz, dx_inv);
}

double grid_sum(const float* grid, int n) {
  double* d = (double*)dmalloc(sizeof(double));
  FS2_HIP(hipMemsetAsync(d, 0, sizeof(double), stream()));
  const int64_t rows = (int64_t)n * n;
  const unsigned nb = (unsigned)(rows < 4096 ? rows : 4096);
  launch("grid_sum", (double)n * n * n * 4.0, k_grid_sum, dim3(nb), dim3(kBlock), 0, grid, n,
         grid_rpitch(n), d);
  double h = 0.0;
  FS2_HIP(hipMemcpyAsync(&h, d, sizeof(double), hipMemcpyDeviceToHost, stream()));
  FS2_HIP(hipStreamSynchronize(stream()));
  dfree(d);
  return h;
}

void power_spectrum(const float* delta_k, int n, double boxsize, int nbins, double dk_bin,
                    double* k, double* P, double* nmodes) {
  if (nbins <= 0) return;
  if (3 * nbins * sizeof(double) > 60000) fail(FS2_ERR_VALUE, "too many P(k) bins: %d", nbins);
  double* d = (double*)dmalloc(sizeof(double) * 3 * nbins);
  try {
    power_spectrum_sums(delta_k, n, boxsize, nbins, dk_bin, d, k, P, nmodes);
  } catch (...) {
    dfree(d);
    throw;
  }
  dfree(d);
}

void power_spectrum_noalloc(const float* delta_k, int n, double boxsize, int nbins,
                            double dk_bin, double* k, double* P, double* nmodes) {
  if (nbins <= 0) return;
  if (nbins > FS2_PK_NOALLOC_MAX_BINS)
    fail(FS2_ERR_VALUE, "%d P(k) bins exceed the fixed scratch of %d bins", nbins,
         FS2_PK_NOALLOC_MAX_BINS);
  int dev = 0;
  FS2_HIP(hipGetDevice(&dev));
  double* d = (dev >= 0 && dev < kMaxDevices) ? g_pk_scratch[dev] : nullptr;
  if (!d) fail(FS2_ERR_DEVICE, "P(k) scratch not set up on device %d", dev);
  power_spectrum_sums(delta_k, n, boxsize, nbins, dk_bin, d, k, P, nmodes);
}

}  // namespace fs2
