// Device-side helpers shared by the HIP kernel files.  Kept out of
// fs2_internal.h so that the host-only sources (fof.cpp, host_numerics.cpp)
// also build with a plain host compiler, e.g. under the sanitizers.
#pragma once

#include <hip/hip_runtime.h>

#include "fs2_internal.h"

namespace fs2 {

// Workgroups are dispatched round-robin over the 8 XCDs of the MI355X; this
// remaps the hardware block index so that consecutive logical blocks run on
// the same XCD and share its L2.
__device__ __forceinline__ int xcd_remap(unsigned bid, unsigned nblocks) {
  const unsigned per = nblocks >> 3, rem = nblocks & 7u;
  const unsigned xcd = bid & 7u, local = bid >> 3;
  return (int)(xcd < rem ? xcd * (per + 1) + local : rem * (per + 1) + (xcd - rem) * per + local);
}

// Green x gradient multiplier f2 = f1 k_a / |k|^2 (py/pm.cpp:460-504) with a
// hardware reciprocal (<= 1 ulp) in place of an IEEE division: this sits in
// the inner loop of the fused force pass, three times per mode.  The
// forward and adjoint kernels share it, so M^T = -M holds for the rounded
// operator too.
__device__ __forceinline__ float green_f2(float f1, float k2, float ka) {
  return f1 * __builtin_amdgcn_rcpf(k2) * ka;
}

// One coordinate of the COLA kick (cola.cpp:79-88) followed by the drift
// (cola.cpp:117-124), in the reference's mixed-precision order: the kick's
// product with -1.5 Omega_m formed in double and rounded to float, v
// updated in float, and the drift's increment formed in double and rounded
// once into x.  Contraction is off, so every kernel that uses it rounds the
// same way.  kick = 0 runs the drift alone (f is not read then).
// KickDrift is declared in fs2_internal.h.
__device__ __forceinline__ float kick_drift_coord(const KickDrift& c, float x, float& v,
                                                  const float* f, float d1, float d2) {
#pragma clang fp contract(off)
  if (c.kick) {
    const float s = *f + d1 * c.q1 + d2 * c.q2;
    const float acc = (float)(c.m15om * (double)s);
    v = v + acc * c.kf;
  }
  const float lpt = d1 * c.da1 + d2 * c.da2;
  const double rhs = (double)v * c.dt + (double)lpt;
  return (float)((double)x + rhs);
}

}  // namespace fs2
