// COLA kick / drift kernels for gfx950 (py/cola.cpp:48-148) and the
// standard leapfrog ones (py/leapfrog.cpp:18-94).
//
// The FP64 time coefficients come from the host; the per-particle update
// keeps the reference's mixed-precision expression order, so contraction is
// disabled here.
#include "fs2_internal.h"
#include "fs2_device.h"

#pragma clang fp contract(off)

namespace fs2 {
namespace {

constexpr int kBlock = 256;
inline unsigned blocks_for(int64_t n) { return (unsigned)((n + kBlock - 1) / kBlock); }

// cola.cpp:79-88:  Float a = -1.5*om*(f + dx1*q1 + dx2*q2)  (double product,
// rounded to float), then v += a*kick_factor in float.
__global__ void k_kick(float* __restrict__ v, const float* __restrict__ f,
                       const float* __restrict__ dx1, const float* __restrict__ dx2,
                       int64_t stride, int64_t np, double m15om, float kf, float q1, float q2) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= np) return;
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const int64_t j = i + a * stride;
    const float s = f[j] + dx1[j] * q1 + dx2[j] * q2;
    const float acc = (float)(m15om * (double)s);
    v[j] = v[j] + acc * kf;
  }
}

// cola.cpp:117-124:  x += v*dt + (dx1*da1 + dx2*da2)  with dt double, so the
// sum is formed in double and rounded once into x.
__global__ void k_drift(float* __restrict__ x, const float* __restrict__ v,
                        const float* __restrict__ dx1, const float* __restrict__ dx2,
                        int64_t stride, int64_t np, KickDrift c) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= np) return;
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const int64_t j = i + a * stride;
    float vj = v[j];
    x[j] = kick_drift_coord(c, x[j], vj, nullptr, dx1[j], dx2[j]);
  }
}

// k_kick followed by k_drift on the same particle, fused: v is written by
// the kick expression and that rounded FP32 value feeds the drift, exactly
// as when the two kernels run in sequence.  One pass reads x v f dx1 dx2
// and writes x v.  The drifted positions go to x, read from xin (the same
// array, or the previous checkpoint of a differentiable run).
__global__ void k_kick_drift(float* x, const float* xin, float* __restrict__ v,
                             const float* __restrict__ f, const float* __restrict__ dx1,
                             const float* __restrict__ dx2, int64_t stride, int64_t np,
                             KickDrift c) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= np) return;
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const int64_t j = i + a * stride;
    float vj = v[j];
    x[j] = kick_drift_coord(c, xin[j], vj, f + j, dx1[j], dx2[j]);
    v[j] = vj;
  }
}

// dst[i] = src[perm[i]]: applies the kd-tree order of the FoF finder to a
// particle plane (32-bit fields) or the id array (64-bit).
template <typename T>
__global__ void k_gather(T* __restrict__ dst, const T* __restrict__ src,
                         const int32_t* __restrict__ perm, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = src[perm[i]];
}

// cola.cpp:141-145: v + Dv dx1 + D2v dx2 (Float coefficients).
__global__ void k_velocity(float* __restrict__ out, const float* __restrict__ v,
                           const float* __restrict__ dx1, const float* __restrict__ dx2,
                           int64_t stride, int64_t np, float Dv, float D2v) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= np) return;
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const int64_t j = i + a * stride;
    out[j] = v[j] + Dv * dx1[j] + D2v * dx2[j];
  }
}

__global__ void k_scale(float* __restrict__ out, const float* __restrict__ in, int64_t stride,
                        int64_t np, float s) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= np) return;
#pragma unroll
  for (int a = 0; a < 3; ++a) out[i + a * stride] = s * in[i + a * stride];
}

// leapfrog.cpp:31-33: v = dx1 Dv + dx2 D2v (Float)
__global__ void k_lf_velocity(float* __restrict__ v, const float* __restrict__ dx1,
                              const float* __restrict__ dx2, int64_t stride, int64_t np, float Dv,
                              float D2v) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= np) return;
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const int64_t j = i + a * stride;
    v[j] = dx1[j] * Dv + dx2[j] * D2v;
  }
}

// leapfrog.cpp:61-65: v += -1.5*om*f*kick_factor, the product in double
// (om is double) and rounded once into v.
__global__ void k_lf_kick(float* __restrict__ v, const float* __restrict__ f, int64_t stride,
                          int64_t np, double m15om, float kf) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= np) return;
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const int64_t j = i + a * stride;
    v[j] = (float)((double)v[j] + m15om * (double)f[j] * (double)kf);
  }
}

// leapfrog.cpp:86-90: x += v*dt with dt double
__global__ void k_lf_drift(float* __restrict__ x, const float* __restrict__ v, int64_t stride,
                           int64_t np, double dt) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= np) return;
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const int64_t j = i + a * stride;
    x[j] = (float)((double)x[j] + (double)v[j] * dt);
  }
}

}  // namespace

void leapfrog_initial_velocity(float* v, const float* dx1, const float* dx2, int64_t stride,
                               int64_t np, float Dv, float D2v) {
  if (np == 0) return;
  launch("leapfrog_velocity", (double)np * 36.0, k_lf_velocity, dim3(blocks_for(np)),
         dim3(kBlock), 0, v, dx1, dx2, stride, np, Dv, D2v);
}

void leapfrog_kick(float* v, const float* f, int64_t stride, int64_t np, double m15om, float kf) {
  if (np == 0) return;
  launch("leapfrog_kick", (double)np * 36.0, k_lf_kick, dim3(blocks_for(np)), dim3(kBlock), 0,
         v, f, stride, np, m15om, kf);
}

void leapfrog_drift(float* x, const float* v, int64_t stride, int64_t np, double dt) {
  if (np == 0) return;
  launch("leapfrog_drift", (double)np * 36.0, k_lf_drift, dim3(blocks_for(np)), dim3(kBlock), 0,
         x, v, stride, np, dt);
}

void cola_kick(float* v, const float* f, const float* dx1, const float* dx2, int64_t stride,
               int64_t np, double m15om, float kf, float q1, float q2) {
  if (np == 0) return;
  launch("cola_kick", (double)np * 60.0, k_kick, dim3(blocks_for(np)), dim3(kBlock), 0, v, f,
         dx1, dx2, stride, np, m15om, kf, q1, q2);
}

void cola_drift(float* x, const float* v, const float* dx1, const float* dx2, int64_t stride,
                int64_t np, double dt, float da1, float da2) {
  if (np == 0) return;
  KickDrift c{};
  c.kick = 0;
  c.dt = dt;
  c.da1 = da1;
  c.da2 = da2;
  launch("cola_drift", (double)np * 60.0, k_drift, dim3(blocks_for(np)), dim3(kBlock), 0, x, v,
         dx1, dx2, stride, np, c);
}

void cola_kick_drift(float* x, const float* xin, float* v, const float* f, const float* dx1,
                     const float* dx2, int64_t stride, int64_t np, double m15om, float kf,
                     float q1, float q2, double dt, float da1, float da2) {
  if (np == 0) return;
  const KickDrift c{1, m15om, kf, q1, q2, dt, da1, da2};
  launch("cola_kick_drift", (double)np * 84.0, k_kick_drift, dim3(blocks_for(np)), dim3(kBlock),
         0, x, xin, v, f, dx1, dx2, stride, np, c);
}

void gather_u32(uint32_t* dst, const uint32_t* src, const int32_t* perm, int64_t n) {
  if (n == 0) return;
  launch("gather_u32", (double)n * 12.0, k_gather<uint32_t>, dim3(blocks_for(n)), dim3(kBlock),
         0, dst, src, perm, n);
}

void gather_u64(uint64_t* dst, const uint64_t* src, const int32_t* perm, int64_t n) {
  if (n == 0) return;
  launch("gather_u64", (double)n * 20.0, k_gather<uint64_t>, dim3(blocks_for(n)), dim3(kBlock),
         0, dst, src, perm, n);
}

void cola_velocity(float* out, const float* v, const float* dx1, const float* dx2,
                   int64_t stride, int64_t np, float Dv, float D2v) {
  if (np == 0) return;
  launch("cola_velocity", (double)np * 48.0, k_velocity, dim3(blocks_for(np)), dim3(kBlock), 0,
         out, v, dx1, dx2, stride, np, Dv, D2v);
}

void scale_field(float* out, const float* in, int64_t stride, int64_t np, float s) {
  if (np == 0) return;
  launch("scale_field", (double)np * 24.0, k_scale, dim3(blocks_for(np)), dim3(kBlock), 0, out,
         in, stride, np, s);
}

}  // namespace fs2
