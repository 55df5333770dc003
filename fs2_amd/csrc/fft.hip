// Hand-written 3D r2c / c2r FFTs for gfx950.
//
// Replaces FFTW3f(-MPI) as used by py/fft.cpp:19-100 (unnormalised, sign -1
// forward, in place on a padded [n][n][2(n/2+1)] real grid).  The transform
// is three HBM passes, each a batch of 1D FFTs staged through LDS:
//
//   forward r2c:  z pass (rows, r2c via an n/2-point complex FFT)
//                 y pencils (stride n/2+1), x pencils (stride n(n/2+1))
//   inverse c2r:  x pencils, y pencils, z pass (c2r)
//
// The inverse runs the complex passes first and the c2r along z last, so the
// imaginary parts of the kz = 0 and kz = n/2 planes are dropped after the x/y
// transforms -- the FFTW / numpy irfftn result for the non-Hermitian planes
// that multiplying by an odd function of k produces.
//
// Inside a pass every pencil is transformed by a Stockham autosort FFT in LDS
// (Govindaraju et al., SC'08 formulation): stage s with radix R and span Ns
// reads v[r] = in[j + r N/R], multiplies by w^{r (j mod Ns)}, w =
// e^{-2 pi i/(Ns R)}, applies a radix-R DFT, and writes
// out[(j/Ns) Ns R + j mod Ns + r Ns].  Radices 2, 3, 4, 8, 16; twiddles come
// from an FP64-computed table rounded once to FP32.  Inverse transforms use
// IDFT(x) = conj(DFT(conj(x))).
#include <algorithm>
#include <cmath>
#include <map>
#include <mutex>
#include <type_traits>

#include "fs2_internal.h"
#include "fs2_device.h"
#include "fft_device.h"

namespace fs2 {
namespace {

using namespace fftdev;

// ---- Green x gradient multiplier (py/pm.cpp:460-504) ---------------------------------------
struct GreenArg {
  int axis;
  int n;
  float f1;
};

__device__ __forceinline__ float2 green_mul(float2 v, int ix, int iy, int iz, const GreenArg& g) {
  const int n = g.n;
  const int kx = ix <= n / 2 ? ix : ix - n;
  const int ky = iy <= n / 2 ? iy : iy - n;
  if (kx == 0 && ky == 0 && iz == 0) return make_float2(0.f, 0.f);
  const float k0 = (float)kx, k1 = (float)ky, k2 = (float)iz;
  const float ka = g.axis == 0 ? k0 : (g.axis == 1 ? k1 : k2);
  const float f2 = green_f2(g.f1, k0 * k0 + k1 * k1 + k2 * k2, ka);
  return make_float2(-f2 * v.y, f2 * v.x);
}

// ---- kernels --------------------------------------------------------------------------------
// Strided complex pencils: pencil (o, c) has elements src[o*outer + c + j*elem].
// A block takes P adjacent columns c of one outer index o and stages the N x P
// tile in LDS as [j][P], so global accesses are P-wide contiguous segments.
template <int N, int P, bool INV, bool GREEN>
__global__ void __launch_bounds__(Plan<N>::T * P)
k_fft_pencil(const float2* src, float2* dst, int ncols, int64_t elem_stride,
             int64_t outer_stride, int nchunks, int nblocks, const float2* __restrict__ tw,
             GreenArg g, int64_t bstride) {
  constexpr int T = Plan<N>::T;
  using L = RowsLayout<P>;
  __shared__ float2 s[L::rows(N) * P];
  __shared__ float2 stw[N];
  TwiddleStage<N, T * P> tws;
  tws.load(tw);   // stored after the pencil's loads; fft_regs' first barrier orders it
  src += blockIdx.y * bstride;   // mesh of this block's box (batched launches)
  dst += blockIdx.y * bstride;
  const int tid = threadIdx.x;
  const int p = tid % P, t = tid / P;
  const int bid = xcd_remap(blockIdx.x, nblocks);
  const int o = bid / nchunks;
  const int c = (bid - o * nchunks) * P + p;
  const bool valid = c < ncols;
  // 32-bit element offsets: grids up to 1024^3 stay below 2^31 float2
  const int base = o * (int)outer_stride + c;
  const int es = (int)elem_stride;
  float2 x[N / T];
  // Loads are unconditional: a column past ncols reads the row padding
  // (rows hold grid_cpitch(n) >= nchunks P elements) and is zeroed after, so
  // that no branch end waits for the loads.
#pragma unroll
  for (int i = 0; i < N / T; ++i) {
    const int j = t + i * T;
    float2 v = src[base + j * es];
    if (!valid) v = make_float2(0.f, 0.f);
    if constexpr (GREEN) v = green_mul(v, j, o, c, g);
    if constexpr (INV) v.y = -v.y;
    x[i] = v;
  }
  tws.store(stw);
  fft_regs<N, L>(x, s + p, t, stw);
#pragma unroll
  for (int i = 0; i < N / T; ++i) {
    float2 v = x[i];
    if constexpr (INV) v.y = -v.y;
    if (valid) dst[base + (t + i * T) * es] = v;
  }
}

// Forward x pass fused with the three Green x gradient inverse x passes
// (py/pm.cpp:431-504): each x pencil of the y-transformed mesh is read once,
// transformed forward in LDS, kept in registers as delta_k, and for each
// axis a multiplied by i f2 k_a / |k|^2 and transformed back into the
// corresponding force mesh.  Saves the delta_k write and its three reads.
template <int N, int P>
__global__ void __launch_bounds__(Plan<N>::T * P)
k_fft_x_green3(const float2* src, float2* dst0, float2* dst1, int ncols,
               int64_t elem_stride, int64_t outer_stride, int nchunks, int nblocks,
               const float2* __restrict__ tw, float f1, int64_t bstride) {
  constexpr int T = Plan<N>::T;
  using L = RowsLayout<P>;
  __shared__ float2 s[L::rows(N) * P];
  __shared__ float2 stw[N];
  TwiddleStage<N, T * P> tws;
  tws.load(tw);   // stored after the pencil's loads; fft_regs' first barrier orders it
  const int64_t boff = blockIdx.y * bstride;   // mesh of this block's box (batched launches)
  src += boff;
  dst0 += boff;
  dst1 += boff;
  const int tid = threadIdx.x;
  const int p = tid % P, t = tid / P;
  const int bid = xcd_remap(blockIdx.x, nblocks);
  const int o = bid / nchunks;
  const int c = (bid - o * nchunks) * P + p;
  const bool valid = c < ncols;
  const int base = o * (int)outer_stride + c;
  const int es = (int)elem_stride;
  float2 dk[N / T];
  // unconditional loads, as in k_fft_pencil
#pragma unroll
  for (int i = 0; i < N / T; ++i) {
    const int j = t + i * T;
    const float2 v = src[base + j * es];
    dk[i] = valid ? v : make_float2(0.f, 0.f);
  }
  tws.store(stw);
  fft_regs<N, L>(dk, s + p, t, stw);
  // Along an x pencil ky and kz are constant, so the y and z force spectra
  // are ky and kz times one transform G = IFFT_x(i f1/|k|^2 delta_k): two
  // inverse transforms instead of three, and two meshes written instead of
  // three.  Pass 0 writes the x force to dst0, pass 1 writes G to dst1.
  const int kyi = o <= N / 2 ? o : o - N;
  const float ky = (float)kyi, kz = (float)c;
#pragma unroll 1
  for (int pass = 0; pass < 2; ++pass) {
    // Opaque copies of the thread coordinates: without them the compiler
    // hoists every per-stage LDS address out of this loop and keeps ~100
    // of them live, halving occupancy.
    int tl = t, pl = p;
    asm volatile("" : "+v"(tl), "+v"(pl));
    float2 x[N / T];
#pragma unroll
    for (int i = 0; i < N / T; ++i) {
      const int j = tl + i * T;
      const int kxi = j <= N / 2 ? j : j - N;
      const float kx = (float)kxi;
      const float k2 = kx * kx + ky * ky + kz * kz;
      // green_f2 with k_a = kx (pass 0) or 1 (pass 1); k = 0 -> 0
      const float f2 = k2 > 0.f ? green_f2(f1, k2, pass == 0 ? kx : 1.f) : 0.f;
      // conj(i f2 dk): the inverse transform runs as a conjugated forward one
      x[i] = make_float2(-f2 * dk[i].y, -(f2 * dk[i].x));
    }
    fft_regs<N, L>(x, s + pl, tl, stw);
    if (pass == 0) {
#pragma unroll
      for (int i = 0; i < N / T; ++i)
        if (valid) dst0[base + (tl + i * T) * es] = make_float2(x[i].x, -x[i].y);
    } else {
      // x holds conj(IFFT(i f1/|k|^2 delta_k)) = conj(i Phi); the y pass
      // k_fft_y_inv_dual turns i Phi into the y and z forces
#pragma unroll
      for (int i = 0; i < N / T; ++i)
        if (valid) dst1[base + (tl + i * T) * es] = make_float2(x[i].x, -x[i].y);
    }
  }
}

// Inverse y pass that splits G = IFFT_x(i f1/|k|^2 delta_k) (the second
// output of k_fft_x_green3) into the y and z forces: along a y pencil kz is
// constant, so F_y = IFFT_y(ky G) and F_z = kz IFFT_y(G).  G is read once;
// F_y replaces it in place and F_z goes to dst2.
template <int N, int P>
__global__ void __launch_bounds__(Plan<N>::T * P)
k_fft_y_inv_dual(float2* g1, float2* dst2, int ncols, int64_t elem_stride, int64_t outer_stride,
                 int nchunks, int nblocks, const float2* __restrict__ tw, int64_t bstride) {
  constexpr int T = Plan<N>::T;
  using L = RowsLayout<P>;
  __shared__ float2 s[L::rows(N) * P];
  __shared__ float2 stw[N];
  TwiddleStage<N, T * P> tws;
  tws.load(tw);   // stored after the pencil's loads; fft_regs' first barrier orders it
  g1 += blockIdx.y * bstride;
  dst2 += blockIdx.y * bstride;
  const int tid = threadIdx.x;
  const int p = tid % P, t = tid / P;
  const int bid = xcd_remap(blockIdx.x, nblocks);
  const int o = bid / nchunks;
  const int c = (bid - o * nchunks) * P + p;
  const bool valid = c < ncols;
  const int base = o * (int)outer_stride + c;
  const int es = (int)elem_stride;
  const float kz = (float)c;
  float2 gk[N / T];
  // unconditional loads, as in k_fft_pencil
#pragma unroll
  for (int i = 0; i < N / T; ++i) {
    const float2 v = g1[base + (t + i * T) * es];
    gk[i] = valid ? v : make_float2(0.f, 0.f);
  }
  tws.store(stw);
#pragma unroll 1
  for (int pass = 0; pass < 2; ++pass) {
    int tl = t, pl = p;
    asm volatile("" : "+v"(tl), "+v"(pl));
    float2 x[N / T];
#pragma unroll
    for (int i = 0; i < N / T; ++i) {
      const int j = tl + i * T;
      const float m = pass == 0 ? (float)(j <= N / 2 ? j : j - N) : 1.f;   // ky or 1
      x[i] = make_float2(m * gk[i].x, -(m * gk[i].y));
    }
    fft_regs<N, L>(x, s + pl, tl, stw);
    if (pass == 0) {
#pragma unroll
      for (int i = 0; i < N / T; ++i)
        if (valid) g1[base + (tl + i * T) * es] = make_float2(x[i].x, -x[i].y);
    } else {
#pragma unroll
      for (int i = 0; i < N / T; ++i)
        if (valid) dst2[base + (tl + i * T) * es] = make_float2(kz * x[i].x, -(kz * x[i].y));
    }
  }
}

// ---- adjoint of the force-mesh operator ------------------------------------
// rho_bar = -sum_a M_a P_a with M_a = c2r . (i f1 k_a/|k|^2) . r2c needs the
// spectrum S = sum_a k_a P_a^.  k_z is constant along y pencils and k_y, k_z
// along x pencils, so S is formed while transforming: a dual forward y pass
// makes Q = k_y Y(P_y) + k_z Y(P_z), a fused x pass forms S = k_x X(P_x) +
// X(Q), multiplies by -i f1/|k|^2 and transforms back along x.

// Q = ky Y(a) + kz Y(b) along y pencils (a, b z-transformed); Q overwrites a.
template <int N, int P>
__global__ void __launch_bounds__(Plan<N>::T * P)
k_fft_y_fwd_dual(float2* a, const float2* b, int ncols, int64_t elem_stride,
                 int64_t outer_stride, int nchunks, int nblocks, const float2* __restrict__ tw) {
  constexpr int T = Plan<N>::T;
  using L = RowsLayout<P>;
  __shared__ float2 s[L::rows(N) * P];
  __shared__ float2 stw[N];
  stage_twiddles<N>(stw, tw);
  const int tid = threadIdx.x;
  const int p = tid % P, t = tid / P;
  const int bid = xcd_remap(blockIdx.x, nblocks);
  const int o = bid / nchunks;
  const int c = (bid - o * nchunks) * P + p;
  const bool valid = c < ncols;
  const int base = o * (int)outer_stride + c;
  const int es = (int)elem_stride;
  const float kz = (float)c;
  float2 ya[N / T];
#pragma unroll
  for (int i = 0; i < N / T; ++i) ya[i] = valid ? a[base + (t + i * T) * es] : make_float2(0.f, 0.f);
  fft_regs<N, L>(ya, s + p, t, stw);
  int tl = t, pl = p;
  asm volatile("" : "+v"(tl), "+v"(pl));
  float2 yb[N / T];
#pragma unroll
  for (int i = 0; i < N / T; ++i)
    yb[i] = valid ? b[base + (tl + i * T) * es] : make_float2(0.f, 0.f);
  fft_regs<N, L>(yb, s + pl, tl, stw);
#pragma unroll
  for (int i = 0; i < N / T; ++i) {
    const int j = tl + i * T;
    const float ky = (float)(j <= N / 2 ? j : j - N);
    if (valid)
      a[base + j * es] = make_float2(ky * ya[i].x + kz * yb[i].x, ky * ya[i].y + kz * yb[i].y);
  }
}

// R = IFFT_x(-i f1/|k|^2 (kx X(a) + X(q))) along x pencils; R overwrites a.
template <int N, int P>
__global__ void __launch_bounds__(Plan<N>::T * P)
k_fft_x_div_green(float2* a, const float2* q, int ncols, int64_t elem_stride,
                  int64_t outer_stride, int nchunks, int nblocks, const float2* __restrict__ tw,
                  float f1) {
  constexpr int T = Plan<N>::T;
  using L = RowsLayout<P>;
  __shared__ float2 s[L::rows(N) * P];
  __shared__ float2 stw[N];
  stage_twiddles<N>(stw, tw);
  const int tid = threadIdx.x;
  const int p = tid % P, t = tid / P;
  const int bid = xcd_remap(blockIdx.x, nblocks);
  const int o = bid / nchunks;
  const int c = (bid - o * nchunks) * P + p;
  const bool valid = c < ncols;
  const int base = o * (int)outer_stride + c;
  const int es = (int)elem_stride;
  const float ky = (float)(o <= N / 2 ? o : o - N), kz = (float)c;
  float2 xa[N / T];
#pragma unroll
  for (int i = 0; i < N / T; ++i) xa[i] = valid ? a[base + (t + i * T) * es] : make_float2(0.f, 0.f);
  fft_regs<N, L>(xa, s + p, t, stw);
  int tl = t, pl = p;
  asm volatile("" : "+v"(tl), "+v"(pl));
  float2 x[N / T];
#pragma unroll
  for (int i = 0; i < N / T; ++i)
    x[i] = valid ? q[base + (tl + i * T) * es] : make_float2(0.f, 0.f);
  fft_regs<N, L>(x, s + pl, tl, stw);
  int tm = tl, pm = pl;
  asm volatile("" : "+v"(tm), "+v"(pm));
#pragma unroll
  for (int i = 0; i < N / T; ++i) {
    const int j = tm + i * T;
    const float kx = (float)(j <= N / 2 ? j : j - N);
    const float k2 = kx * kx + ky * ky + kz * kz;
    const float g = k2 > 0.f ? green_f2(f1, k2, 1.f) : 0.f;
    const float2 S = make_float2(kx * xa[i].x + x[i].x, kx * xa[i].y + x[i].y);
    // conj(-i g S) = (g S.y, g S.x): the inverse runs as a conjugated forward FFT
    x[i] = make_float2(g * S.y, g * S.x);
  }
  fft_regs<N, L>(x, s + pm, tm, stw);
#pragma unroll
  for (int i = 0; i < N / T; ++i)
    if (valid) a[base + (tm + i * T) * es] = make_float2(x[i].x, -x[i].y);
}

// z pass, r2c: each row of n = 2M reals (padded to M+1 complex) is read as M
// complex z[m] = x[2m] + i x[2m+1], transformed, and split:
//   X[k] = E - i W^k O,  E = (Z[k] + conj Z[M-k])/2,  O = (Z[k] - conj Z[M-k])/2,
//   W = e^{-2 pi i/n}, k = 0..M.
template <int M, int P>
__global__ void __launch_bounds__(Plan<M>::T * P)
k_fft_z_r2c(float2* data, int64_t nrows, int cpitch, const float2* __restrict__ twM,
            const float2* __restrict__ twN) {
  constexpr int T = Plan<M>::T;
  using L = ContigLayout;
  constexpr int LS = L::size(M);
  __shared__ float2 s[P * LS];
  const int tid = threadIdx.x;
  const int t = tid % T, p = tid / T;
  const int64_t row = (int64_t)blockIdx.x * P + p;
  const bool valid = row < nrows;
  // a row past the end loads the last row (and stores nothing), so that
  // the loads are unconditional and no branch end waits for them
  float2* r = data + (valid ? row : nrows - 1) * cpitch;
  float2* sp = s + p * LS;
  float2 x[M / T];
#pragma unroll
  for (int i = 0; i < M / T; ++i) {
    const int m = t + i * T;
    const float2 v = r[m];
    x[i] = valid ? v : make_float2(0.f, 0.f);
  }
  using S = RowSync<M>;
  fft_regs<M, L, S>(x, sp, t, twM);
  S::sync();   // every thread of the row is done reading the last stage's input
#pragma unroll
  for (int i = 0; i < M / T; ++i) sp[L::at(t + i * T)] = x[i];
  S::sync();
#pragma unroll
  for (int i = 0; i < M / T; ++i) {
    const int k = t + i * T;
    const float2 X = r2c_split(sp[L::at(k)], sp[L::at(k == 0 ? 0 : M - k)], twN[k]);
    if (valid) r[k] = X;
  }
  if (t == 0 && valid) {
    const float2 A = sp[L::at(0)];
    r[M] = make_float2(A.x - A.y, 0.f);
  }
}

// The same r2c with the M-point and split twiddles staged in LDS once per
// block of P rows (for large M, where every one-row block would otherwise
// fetch ~20 twiddles per thread from L2).
template <int M, int P>
__global__ void __launch_bounds__(Plan<M>::T * P)
k_fft_z_r2c_lt(float2* data, int64_t nrows, int cpitch, const float2* __restrict__ twM,
               const float2* __restrict__ twN) {
  constexpr int T = Plan<M>::T;
  using L = ContigLayout;
  constexpr int LS = L::size(M);
  __shared__ float2 s[P * LS];
  __shared__ float2 stM[M];
  __shared__ float2 stN[M];
  for (int i = threadIdx.x; i < M; i += blockDim.x) {
    stM[i] = twM[i];
    stN[i] = twN[i];
  }
  const int tid = threadIdx.x;
  const int t = tid % T, p = tid / T;
  const int64_t row = (int64_t)blockIdx.x * P + p;
  const bool valid = row < nrows;
  // a row past the end loads the last row (and stores nothing), so that
  // the loads are unconditional and no branch end waits for them
  float2* r = data + (valid ? row : nrows - 1) * cpitch;
  float2* sp = s + p * LS;
  float2 x[M / T];
#pragma unroll
  for (int i = 0; i < M / T; ++i) {
    const int m = t + i * T;
    const float2 v = r[m];
    x[i] = valid ? v : make_float2(0.f, 0.f);
  }
  using S = RowSync<M>;
  __syncthreads();   // the twiddles are staged
  fft_regs<M, L, S>(x, sp, t, stM);
  S::sync();   // every thread of the row is done reading the last stage's input
#pragma unroll
  for (int i = 0; i < M / T; ++i) sp[L::at(t + i * T)] = x[i];
  S::sync();
#pragma unroll
  for (int i = 0; i < M / T; ++i) {
    const int k = t + i * T;
    const float2 X = r2c_split(sp[L::at(k)], sp[L::at(k == 0 ? 0 : M - k)], stN[k]);
    if (valid) r[k] = X;
  }
  if (t == 0 && valid) {
    const float2 A = sp[L::at(0)];
    r[M] = make_float2(A.x - A.y, 0.f);
  }
}

// z pass, c2r: the inverse of the split above, with the imaginary parts of
// X[0] and X[M] dropped:  Z[k] = E' + i O',  E' = X[k] + conj X[M-k],
// O' = (X[k] - conj X[M-k]) W^{-k};  z = IDFT_M(Z) = x_even + i x_odd.
template <int M, int P>
__global__ void __launch_bounds__(Plan<M>::T * P)
k_fft_z_c2r(float2* data, int64_t nrows, int cpitch, const float2* __restrict__ twM,
            const float2* __restrict__ twN) {
  constexpr int T = Plan<M>::T;
  using L = ContigLayout;
  constexpr int LS = L::size(M);
  __shared__ float2 s[P * LS];
  const int tid = threadIdx.x;
  const int t = tid % T, p = tid / T;
  const int64_t row = (int64_t)blockIdx.x * P + p;
  const bool valid = row < nrows;
  // unconditional loads, as in k_fft_z_r2c
  float2* r = data + (valid ? row : nrows - 1) * cpitch;
  float2* sp = s + p * LS;
  // X[k] and X[M - k] straight from the row (the mirrored loads hit the
  // lines the direct ones bring in)
  float2 z[M / T], zb[M / T];
#pragma unroll
  for (int i = 0; i < M / T; ++i) {   // every load issued before any use
    const int k = t + i * T;
    z[i] = r[k];
    zb[i] = r[M - k];
  }
#pragma unroll
  for (int i = 0; i < M / T; ++i) {
    const int k = t + i * T;
    const float2 A = valid ? z[i] : make_float2(0.f, 0.f);
    const float2 B = valid ? zb[i] : make_float2(0.f, 0.f);
    z[i] = c2r_merge(A, B, twN[k], k == 0);
  }
  fft_regs<M, L, RowSync<M>>(z, sp, t, twM);
#pragma unroll
  for (int i = 0; i < M / T; ++i) {
    const int m = t + i * T;
    if (valid) r[m] = cconj(z[i]);
  }
}

// Persistent c2r z pass for large meshes: each block stages the M-point and
// split twiddles in LDS once and then walks rows P at a time (grid-stride),
// instead of every one-row block fetching ~20 twiddles from L2.

template <int M, int P>
__global__ void __launch_bounds__(Plan<M>::T * P)
k_fft_z_c2r_p(float2* data, int64_t nrows, int cpitch, const float2* __restrict__ twM,
              const float2* __restrict__ twN) {
  constexpr int T = Plan<M>::T;
  using L = ContigLayout;
  constexpr int LS = L::size(M);
  __shared__ float2 s[P * LS];
  __shared__ float2 stM[M];
  __shared__ float2 stN[M];
  for (int i = threadIdx.x; i < M; i += blockDim.x) {
    stM[i] = twM[i];
    stN[i] = twN[i];
  }
  const int tid = threadIdx.x;
  using S = RowSync<M>;
  __syncthreads();   // the twiddles are staged
  for (int64_t row0 = (int64_t)blockIdx.x * P; row0 < nrows; row0 += (int64_t)gridDim.x * P) {
    // opaque thread coordinates: keeps the compiler from holding every LDS
    // address of the transform live across iterations
    int t = tid % T, p = tid / T;
    asm volatile("" : "+v"(t), "+v"(p));
    float2* sp = s + p * LS;
    const int64_t row = row0 + p;
    const bool valid = row < nrows;
    float2* r = data + row * cpitch;
    // X[k] and X[M - k] straight from the row (the mirrored loads hit the
    // lines the direct ones bring in); the transform's first barrier orders
    // its LDS writes after the previous row's last reads
    float2 z[M / T];
#pragma unroll
    for (int i = 0; i < M / T; ++i) {
      const int k = t + i * T;
      const float2 A = valid ? r[k] : make_float2(0.f, 0.f);
      const float2 B = valid ? r[M - k] : make_float2(0.f, 0.f);
      z[i] = c2r_merge(A, B, stN[k], k == 0);
    }
    fft_regs<M, L, S>(z, sp, t, stM);
#pragma unroll
    for (int i = 0; i < M / T; ++i)
      if (valid) r[t + i * T] = cconj(z[i]);
  }
}

// ---- host side ---------------------------------------------------------------------------------
const float2* twiddles(int N) {
  static std::mutex mu;
  static std::map<int, float2*> tables;
  std::lock_guard<std::mutex> lock(mu);
  auto it = tables.find(N);
  if (it != tables.end()) return it->second;
  std::vector<float2> h(N);
  for (int k = 0; k < N; ++k) {
    const double a = -2.0 * M_PI * (double)k / (double)N;
    h[k] = make_float2((float)std::cos(a), (float)std::sin(a));
  }
  float2* d = (float2*)dmalloc(sizeof(float2) * N);
  FS2_HIP(hipMemcpy(d, h.data(), sizeof(float2) * N, hipMemcpyHostToDevice));
  tables[N] = d;
  return d;
}

// Columns per block for the strided passes, and rows per block for the z pass.
template <int N> constexpr int pencil_cols() {
  return ((N + N / 16) * 16 * 8 <= 81920 && Plan<N>::T * 16 <= 1024) ? 16 : 8;
}
// rows per block of the LDS-twiddle r2c z pass (M >= 512; at M = 256 the
// register-twiddle kernel with 16 rows per block measured 0.21 ms against 0.30)
constexpr int kZR2CRows = 2;
template <int M> constexpr int zrows() {   // one-wave blocks from M = 512 on (measured)
  return Plan<M>::T >= 256 ? 1 : (M >= 512 ? 64 : 256) / Plan<M>::T;
}

template <int n>
struct Impl {
  static constexpr int N = n;
  static constexpr int M = n / 2;
  static constexpr int P = pencil_cols<N>();
  static constexpr int PZ = zrows<M>();

  static double grid_bytes() { return (double)n * n * (n / 2 + 1) * 8.0; }

  // nb meshes of grid_floats(n) floats each, back to back: the z passes
  // treat them as nb*n*n rows, the pencil passes take the box from blockIdx.y.
  static int64_t bstride() { return (int64_t)grid_floats(n) / 2; }   // in float2

  template <bool INV, bool GREEN>
  static void pencil(const char* name, const float2* src, float2* dst, int64_t elem_stride,
                     int64_t outer_stride, GreenArg g, int nb) {
    const int nh = n / 2 + 1;
    const int nchunks = (nh + P - 1) / P;
    const int nblocks = n * nchunks;
    launch(name, 2.0 * grid_bytes() * nb, k_fft_pencil<N, P, INV, GREEN>, dim3(nblocks, nb),
           dim3(Plan<N>::T * P), 0, src, dst, nh, elem_stride, outer_stride, nchunks, nblocks,
           twiddles(N), g, bstride());
  }

  // rows per block of the persistent c2r z pass (M >= 512; at M = 256 the
  // one-shot kernel measured 0.36 ms against 0.38) and its grid
  static constexpr int PP = Plan<M>::T >= 256 ? 1 : 256 / Plan<M>::T;
  // all blocks resident at once, so the grid-stride loop has no second wave
  template <typename K>
  static unsigned persistent_grid(K kernel, int64_t nrows) {
    static int per_cu = 0;
    if (per_cu == 0) {
      int v = 0;
      FS2_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&v, kernel, Plan<M>::T * PP, 0));
      per_cu = v > 0 ? v : 1;
    }
    const int64_t want = (nrows + PP - 1) / PP;
    return (unsigned)std::min<int64_t>(want, (int64_t)device_cus() * per_cu);
  }

  static void z_r2c(float* grid, int nb) {
    const int64_t nrows = (int64_t)n * n * nb;
    const double bytes = ((double)n * n * n * 4.0 + grid_bytes()) * nb;
    // (a persistent r2c, like the c2r below, measured slower at 1024^3)
    if constexpr (M >= 512) {
      constexpr int PL = kZR2CRows;
      launch("fft_z_r2c", bytes, k_fft_z_r2c_lt<M, PL>, dim3((unsigned)((nrows + PL - 1) / PL)),
             dim3(Plan<M>::T * PL), 0, (float2*)grid, nrows, grid_cpitch(n), twiddles(M),
             twiddles(N));
    } else {
      launch("fft_z_r2c", bytes, k_fft_z_r2c<M, PZ>, dim3((unsigned)((nrows + PZ - 1) / PZ)),
             dim3(Plan<M>::T * PZ), 0, (float2*)grid, nrows, grid_cpitch(n), twiddles(M),
             twiddles(N));
    }
  }

  static void z_c2r(float* grid, int nb) {
    const int64_t nrows = (int64_t)n * n * nb;
    const double bytes = ((double)n * n * n * 4.0 + grid_bytes()) * nb;
    if constexpr (M >= 512) {
      launch("fft_z_c2r", bytes, k_fft_z_c2r_p<M, PP>,
             dim3(persistent_grid(k_fft_z_c2r_p<M, PP>, nrows)),
             dim3(Plan<M>::T * PP), 0, (float2*)grid, nrows, grid_cpitch(n), twiddles(M),
             twiddles(N));
    } else {
      launch("fft_z_c2r", bytes, k_fft_z_c2r<M, PZ>, dim3((unsigned)((nrows + PZ - 1) / PZ)),
             dim3(Plan<M>::T * PZ), 0, (float2*)grid, nrows, grid_cpitch(n), twiddles(M),
             twiddles(N));
    }
  }

  static void r2c(float* grid, int nb) {
    const int64_t nh = grid_cpitch(n);
    float2* c = (float2*)grid;
    GreenArg g{0, n, 0.f};
    z_r2c(grid, nb);
    pencil<false, false>("fft_y_fwd", c, c, nh, (int64_t)n * nh, g, nb);
    pencil<false, false>("fft_x_fwd", c, c, (int64_t)n * nh, nh, g, nb);
  }

  static void c2r(float* grid, int nb) {
    const int64_t nh = grid_cpitch(n);
    float2* c = (float2*)grid;
    GreenArg g{0, n, 0.f};
    pencil<true, false>("fft_x_inv", c, c, (int64_t)n * nh, nh, g, nb);
    pencil<true, false>("fft_y_inv", c, c, nh, (int64_t)n * nh, g, nb);
    z_c2r(grid, nb);
  }

  // density -> three real force meshes (pm.cpp:431-508)
  static void force_meshes(float* mesh, float* const out[3], float f1, int nb) {
    force_spectra(mesh, out, f1, nb);
    for (int a = 0; a < 3; ++a) z_c2r(out[a], nb);
  }

  // force_meshes up to the z c2r: out[a] left holding the half-complex
  // spectra after the x and y inverse passes
  static void force_spectra(float* mesh, float* const out[3], float f1, int nb,
                            bool z_done = false) {
    const int64_t nh = grid_cpitch(n);
    float2* c = (float2*)mesh;
    GreenArg g{0, n, 0.f};
    if (!z_done) z_r2c(mesh, nb);
    pencil<false, false>("fft_y_fwd", c, c, nh, (int64_t)n * nh, g, nb);
    {
      const int ncols = n / 2 + 1;
      const int nchunks = (ncols + P - 1) / P;
      const int nblocks = n * nchunks;
      launch("fft_x_fwd_green3", 3.0 * grid_bytes() * nb, k_fft_x_green3<N, P>,
             dim3(nblocks, nb), dim3(Plan<N>::T * P), 0, (const float2*)c, (float2*)out[0],
             (float2*)out[1], ncols, (int64_t)n * nh, nh, nchunks, nblocks, twiddles(N), f1,
             bstride());
    }
    pencil<true, false>("fft_y_inv", (float2*)out[0], (float2*)out[0], nh, (int64_t)n * nh, g, nb);
    {
      const int ncols = n / 2 + 1;
      const int nchunks = (ncols + P - 1) / P;
      const int nblocks = n * nchunks;
      launch("fft_y_inv_dual", 3.0 * grid_bytes() * nb, k_fft_y_inv_dual<N, P>,
             dim3(nblocks, nb), dim3(Plan<N>::T * P), 0, (float2*)out[1], (float2*)out[2],
             ncols, nh, (int64_t)n * nh, nchunks, nblocks, twiddles(N), bstride());
    }
  }

  // p[0..2]: real meshes P_a on entry; p[0] holds rho_bar = -sum_a M_a P_a
  // on return, p[1] and p[2] are clobbered.
  static void adjoint_green(float* const pm[3], float f1) {
    const int64_t nh = grid_cpitch(n);
    const GreenArg g{0, n, 0.f};
    for (int a = 0; a < 3; ++a) z_r2c(pm[a], 1);
    pencil<false, false>("fft_y_fwd", (float2*)pm[0], (float2*)pm[0], nh, (int64_t)n * nh, g, 1);
    const int ncols = n / 2 + 1;
    const int nchunks = (ncols + P - 1) / P;
    const int nblocks = n * nchunks;
    launch("fft_y_fwd_dual", 3.0 * grid_bytes(), k_fft_y_fwd_dual<N, P>, dim3(nblocks),
           dim3(Plan<N>::T * P), 0, (float2*)pm[1], (const float2*)pm[2], ncols, nh,
           (int64_t)n * nh, nchunks, nblocks, twiddles(N));
    launch("fft_x_div_green", 3.0 * grid_bytes(), k_fft_x_div_green<N, P>, dim3(nblocks),
           dim3(Plan<N>::T * P), 0, (float2*)pm[0], (const float2*)pm[1], ncols,
           (int64_t)n * nh, nh, nchunks, nblocks, twiddles(N), f1);
    pencil<true, false>("fft_y_inv", (float2*)pm[0], (float2*)pm[0], nh, (int64_t)n * nh, g, 1);
    z_c2r(pm[0], 1);
  }

  static void green_c2r(const float* dk, float* dst, int axis, float f1) {
    const int64_t nh = grid_cpitch(n);
    float2* c = (float2*)dst;
    GreenArg g{axis, n, f1};
    pencil<true, true>("fft_x_inv_green", (const float2*)dk, c, (int64_t)n * nh, nh, g, 1);
    pencil<true, false>("fft_y_inv", c, c, nh, (int64_t)n * nh, g, 1);
    z_c2r(dst, 1);
  }
};

template <typename F>
void dispatch(int n, F&& f) {
  switch (n) {
    case 4: f(Impl<4>{}); break;
    case 8: f(Impl<8>{}); break;
    case 16: f(Impl<16>{}); break;
    case 32: f(Impl<32>{}); break;
    case 48: f(Impl<48>{}); break;
    case 64: f(Impl<64>{}); break;
    case 96: f(Impl<96>{}); break;
    case 128: f(Impl<128>{}); break;
    case 256: f(Impl<256>{}); break;
    case 512: f(Impl<512>{}); break;
    case 1024: f(Impl<1024>{}); break;
    default: fail(FS2_ERR_VALUE, "FFT size %d not supported: n must be even with prime factors "
                  "2, 3, 5, 7 (4 <= n <= 4096)", n);
  }
}


// ---- sizes without a tuned plan --------------------------------------------------
// Every other even n whose prime factors are 2, 3, 5 and 7 (FFTW takes any
// n; odd n and larger primes are refused): runtime-planned mixed-radix
// Stockham transforms in LDS, one kernel per axis, all offsets 64-bit (n up
// to 4096; a 2048^3 mesh holds 8.7e9 floats).  A workgroup takes P lines
// (P N <= 4096 values, two LDS buffers of 32 KB).  Stage s of radix R with
// Ns the product of the earlier radices maps element j + r N/R (r < R) of
// the source to (j / Ns) Ns R + j % Ns + p Ns of the destination after the
// twiddles W_N^(r (j % Ns) N / (Ns R)) and a length-R DFT whose factors are
// also read from the N-entry FP64-built table.  The real transforms along z
// run an n/2-point complex transform of the packed row with a split (r2c)
// or merge (c2r) pass; the Green x gradient multiplier is a separate pass.
// Results equal numpy's rfftn / irfftn conventions, as the tuned passes.
constexpr int kGenThreads = 256;
constexpr int kGenValues = 4096;
constexpr int kGenMaxStages = 24;

struct GenPlan {
  int nst;
  int R[kGenMaxStages];
};

template <int R>
__device__ __forceinline__ void gen_stage(const float2* a, float2* b, int N, int P, int Ns,
                                          const float2* __restrict__ tw, bool inv) {
  const int NR = N / R;
  const int tstep = N / (Ns * R);
  for (int q = threadIdx.x; q < P * NR; q += blockDim.x) {
    const int c = q / NR, j = q - c * NR;
    const int k = j % Ns;
    float2 v[R];
#pragma unroll
    for (int r = 0; r < R; ++r) v[r] = a[c * N + j + r * NR];
#pragma unroll
    for (int r = 1; r < R; ++r) {
      float2 w = tw[(int)(((int64_t)r * k * tstep) % N)];
      if (inv) w.y = -w.y;
      v[r] = cmul(v[r], w);
    }
    const int idx = (j / Ns) * Ns * R + k;
#pragma unroll
    for (int p = 0; p < R; ++p) {
      float2 acc = v[0];
#pragma unroll
      for (int r = 1; r < R; ++r) {
        float2 w = tw[((p * r) % R) * NR];
        if (inv) w.y = -w.y;
        acc = cadd(acc, cmul(v[r], w));
      }
      b[c * N + idx + p * Ns] = acc;
    }
  }
}

// Lines (o, c) of nb meshes: element j at data + box*bstride + o*ostride +
// c*cstride + j*elem; the grid is (nouter * nchunks, nb), chunk = P lines.
__global__ void __launch_bounds__(kGenThreads)
k_gen_lines(float2* __restrict__ data, int N, int P, int64_t elem, int64_t ostride,
            int64_t cstride, int ncols, int nchunks, int64_t bstride,
            const float2* __restrict__ tw, GenPlan plan, int inv) {
  extern __shared__ float2 gls[];
  float2* a = gls;
  float2* b = gls + P * N;
  const int o = blockIdx.x / nchunks, ch = blockIdx.x - o * nchunks;
  const int c0 = ch * P;
  const int pc = min(P, ncols - c0);
  float2* d = data + (int64_t)blockIdx.y * bstride + (int64_t)o * ostride + (int64_t)c0 * cstride;
  const bool rows = elem == 1;   // contiguous lines: consecutive threads along j
  for (int i = threadIdx.x; i < pc * N; i += blockDim.x) {
    const int c = rows ? i / N : i % pc, j = rows ? i - (i / N) * N : i / pc;
    a[c * N + j] = d[(int64_t)c * cstride + (int64_t)j * elem];
  }
  __syncthreads();
  int Ns = 1;
  for (int s = 0; s < plan.nst; ++s) {
    const int R = plan.R[s];
    switch (R) {
      case 2: gen_stage<2>(a, b, N, pc, Ns, tw, inv); break;
      case 3: gen_stage<3>(a, b, N, pc, Ns, tw, inv); break;
      case 4: gen_stage<4>(a, b, N, pc, Ns, tw, inv); break;
      case 5: gen_stage<5>(a, b, N, pc, Ns, tw, inv); break;
      case 7: gen_stage<7>(a, b, N, pc, Ns, tw, inv); break;
      default: gen_stage<8>(a, b, N, pc, Ns, tw, inv); break;
    }
    __syncthreads();
    float2* t = a;
    a = b;
    b = t;
    Ns *= R;
  }
  for (int i = threadIdx.x; i < pc * N; i += blockDim.x) {
    const int c = rows ? i / N : i % pc, j = rows ? i - (i / N) * N : i / pc;
    d[(int64_t)c * cstride + (int64_t)j * elem] = a[c * N + j];
  }
}

// r2c split of rows holding Z = FFT_M(x[2m] + i x[2m+1]) in their first M
// values: X[k] = (Z_k + conj Z_{M-k})/2 - (i/2) W_N^k (Z_k - conj Z_{M-k}),
// k = 0..M, in place; one work item per pair (k, M-k).
__global__ void k_gen_split(float2* __restrict__ data, int64_t nrows, int M, int cpitch,
                            const float2* __restrict__ twN) {
  const int64_t per = M / 2 + 1;
  for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < nrows * per;
       q += (int64_t)gridDim.x * blockDim.x) {
  const int64_t row = q / per;
  const int k = (int)(q - row * per);
  float2* d = data + row * cpitch;
  if (k == 0) {
    const float2 z = d[0];
    d[0] = make_float2(z.x + z.y, 0.f);
    d[M] = make_float2(z.x - z.y, 0.f);
    continue;
  }
  const int kk = M - k;
  const float2 zk = d[k], zm = d[kk];
  auto one = [&](float2 a, float2 b, int kw) {   // (a + conj b)/2 - (i/2) W^kw (a - conj b)
    const float2 s = make_float2(0.5f * (a.x + b.x), 0.5f * (a.y - b.y));
    const float2 t = cmul(twN[kw], make_float2(0.5f * (a.x - b.x), 0.5f * (a.y + b.y)));
    return make_float2(s.x + t.y, s.y - t.x);
  };
  const float2 xk = one(zk, zm, k);
  if (kk != k) d[kk] = one(zm, zk, kk);
  d[k] = xk;
  }
}

// c2r merge: Z_k = (X_k + conj X_{M-k}) + i conj(W_N^k) (X_k - conj X_{M-k})
// for k < M, the imaginary parts of X_0 and X_M dropped (numpy's irfft),
// so that the inverse M-point transform gives x[2m] + i x[2m+1].
__global__ void k_gen_merge(float2* __restrict__ data, int64_t nrows, int M, int cpitch,
                            const float2* __restrict__ twN) {
  const int64_t per = M / 2 + 1;
  for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < nrows * per;
       q += (int64_t)gridDim.x * blockDim.x) {
  const int64_t row = q / per;
  const int k = (int)(q - row * per);
  float2* d = data + row * cpitch;
  if (k == 0) {
    const float a = d[0].x, b = d[M].x;
    d[0] = make_float2(a + b, a - b);
    continue;
  }
  const int kk = M - k;
  const float2 xk = d[k], xm = d[kk];
  auto one = [&](float2 a, float2 b, int kw) {
    const float2 e = make_float2(a.x + b.x, a.y - b.y);
    const float2 w = cconj(twN[kw]);
    const float2 o = cmul(make_float2(a.x - b.x, a.y + b.y), w);
    return make_float2(e.x - o.y, e.y + o.x);
  };
  const float2 zk = one(xk, xm, k);
  if (kk != k) d[kk] = one(xm, xk, kk);
  d[k] = zk;
  }
}

// The elementwise generic kernels loop grid-stride over at most this many
// workgroups: a 2048^3 mesh has more half-complex values than a launch may
// have threads (2^32).
constexpr int64_t kGenMaxBlocks = 1 << 20;
inline unsigned gen_blocks(int64_t items) {
  return (unsigned)std::min<int64_t>((items + 255) / 256, kGenMaxBlocks);
}

// dst = i f2_axis(k) src over the half-complex grid (green_mul); for
// axis < 0: dst = -sum_a i f2_a src_a (the adjoint sum) of src0..2.
__global__ void k_gen_green(const float2* __restrict__ s0, const float2* __restrict__ s1,
                            const float2* __restrict__ s2, float2* __restrict__ dst, int n,
                            int cpitch, int axis, float f1, int64_t bstride) {
  const int nh = n / 2 + 1;
  const int64_t total = (int64_t)n * n * nh;
  for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < total;
       q += (int64_t)gridDim.x * blockDim.x) {
  const int iz = (int)(q % nh);
  const int64_t rxy = q / nh;
  const int iy = (int)(rxy % n), ix = (int)(rxy / n);
  const int64_t at = (int64_t)blockIdx.y * bstride + rxy * cpitch + iz;
  if (axis >= 0) {
    dst[at] = green_mul(s0[at], ix, iy, iz, GreenArg{axis, n, f1});
  } else {
    const float2 a = green_mul(s0[at], ix, iy, iz, GreenArg{0, n, f1});
    const float2 b = green_mul(s1[at], ix, iy, iz, GreenArg{1, n, f1});
    const float2 c = green_mul(s2[at], ix, iy, iz, GreenArg{2, n, f1});
    dst[at] = make_float2(-(a.x + b.x + c.x), -(a.y + b.y + c.y));
  }
  }
}

bool gen_factors_ok(int m) {
  if (m < 1) return false;
  for (int p : {2, 3, 5, 7})
    while (m % p == 0) m /= p;
  return m == 1;
}

GenPlan gen_plan(int N) {
  GenPlan pl{};
  int m = N;
  for (int r : {8, 4, 2, 3, 5, 7}) {
    while (m % r == 0) {
      if (pl.nst >= kGenMaxStages) fail(FS2_ERR_VALUE, "FFT plan of %d too long", N);
      pl.R[pl.nst++] = r;
      m /= r;
    }
  }
  if (m != 1) fail(FS2_ERR_VALUE, "FFT size %d has a prime factor above 7", N);
  return pl;
}

struct Gen {
  static bool supported(int n) { return n >= 4 && n <= 4096 && n % 2 == 0 && gen_factors_ok(n); }
  static int64_t bstride(int n) { return (int64_t)grid_floats(n) / 2; }
  // complex transforms of length N along lines (o < nouter, c < ncols)
  static void lines(const char* name, float2* data, int N, int64_t elem, int nouter,
                    int64_t ostride, int ncols, int64_t cstride, int nb, bool inv, int n) {
    const int P = std::max(1, std::min(16, kGenValues / N));
    const int nchunks = (ncols + P - 1) / P;
    const double bytes = 16.0 * (double)N * nouter * ncols * nb;
    launch(name, bytes, k_gen_lines, dim3((unsigned)((int64_t)nouter * nchunks), nb),
           dim3(kGenThreads), (size_t)2 * P * N * sizeof(float2), data, N, P, elem, ostride,
           cstride, ncols, nchunks, bstride(n), twiddles(N), gen_plan(N), inv ? 1 : 0);
  }
  static void y_pass(float2* c, int n, int nb, bool inv) {
    const int64_t cp = grid_cpitch(n);
    lines(inv ? "fft_y_inv" : "fft_y_fwd", c, n, cp, n, (int64_t)n * cp, n / 2 + 1, 1, nb, inv, n);
  }
  static void x_pass(float2* c, int n, int nb, bool inv) {
    const int64_t cp = grid_cpitch(n);
    lines(inv ? "fft_x_inv" : "fft_x_fwd", c, n, (int64_t)n * cp, n, cp, n / 2 + 1, 1, nb, inv,
          n);
  }
  static void z_r2c(float* grid, int n, int nb) {
    const int M = n / 2;
    const int64_t cp = grid_cpitch(n);
    float2* c = (float2*)grid;
    lines("fft_z_r2c", c, M, 1, n * nb, (int64_t)n * cp, n, cp, 1, false, n);
    const int64_t nrows = (int64_t)n * n * nb, items = nrows * (M / 2 + 1);
    launch("fft_z_split", 16.0 * (double)nrows * (M + 1), k_gen_split, dim3(gen_blocks(items)),
           dim3(256), 0, c, nrows, M, (int)cp, twiddles(n));
  }
  static void z_c2r(float* grid, int n, int nb) {
    const int M = n / 2;
    const int64_t cp = grid_cpitch(n);
    float2* c = (float2*)grid;
    const int64_t nrows = (int64_t)n * n * nb, items = nrows * (M / 2 + 1);
    launch("fft_z_merge", 16.0 * (double)nrows * (M + 1), k_gen_merge, dim3(gen_blocks(items)),
           dim3(256), 0, c, nrows, M, (int)cp, twiddles(n));
    lines("fft_z_c2r", c, M, 1, n * nb, (int64_t)n * cp, n, cp, 1, true, n);
  }
  static void r2c(float* grid, int n, int nb) {
    z_r2c(grid, n, nb);
    y_pass((float2*)grid, n, nb, false);
    x_pass((float2*)grid, n, nb, false);
  }
  static void c2r(float* grid, int n, int nb) {
    x_pass((float2*)grid, n, nb, true);
    y_pass((float2*)grid, n, nb, true);
    z_c2r(grid, n, nb);
  }
  static void green(const float* s0, const float* s1, const float* s2, float* dst, int n,
                    int axis, float f1, int nb) {
    const int64_t total = (int64_t)n * n * (n / 2 + 1);
    launch("fft_green", 16.0 * (double)total * nb, k_gen_green, dim3(gen_blocks(total), nb),
           dim3(256), 0, (const float2*)s0,
           (const float2*)s1, (const float2*)s2, (float2*)dst, n, grid_cpitch(n), axis, f1,
           bstride(n));
  }
  static void force_meshes(float* mesh, float* const out[3], int n, float f1, int nb) {
    r2c(mesh, n, nb);
    for (int a = 0; a < 3; ++a) {
      green(mesh, nullptr, nullptr, out[a], n, a, f1, nb);
      c2r(out[a], n, nb);
    }
  }
  static void adjoint_green(float* const p[3], int n, float f1) {
    for (int a = 0; a < 3; ++a) r2c(p[a], n, 1);
    green(p[0], p[1], p[2], p[0], n, -1, f1, 1);
    c2r(p[0], n, 1);
  }
  static void green_c2r(const float* dk, float* dst, int n, int axis, float f1) {
    green(dk, nullptr, nullptr, dst, n, axis, f1, 1);
    c2r(dst, n, 1);
  }
};

bool tuned(int n) {
  switch (n) {
    case 4: case 8: case 16: case 32: case 48: case 64: case 96: case 128: case 256:
    case 512: case 1024:
      return true;
    default:
      return false;
  }
}
}  // namespace

static_assert(1024LL * 1024 * (((1024 / 2 + 1) + 15) & ~15) < (1LL << 31),
              "32-bit float2 offsets in the pencil kernels");

bool fft_supported(int n) { return tuned(n) || Gen::supported(n); }

void fft_prepare(int n) {
  if (tuned(n)) {
    dispatch(n, [&](auto impl) {
      twiddles(decltype(impl)::N);
      twiddles(decltype(impl)::M);
    });
  } else if (Gen::supported(n)) {
    twiddles(n);
    twiddles(n / 2);
  } else {
    fail(FS2_ERR_VALUE, "FFT size %d not supported: n must be even with prime factors 2, 3, 5, 7 "
         "(4 <= n <= 4096)", n);
  }
}

void fft_r2c_3d(float* grid, int n, int nb) {
  if (!tuned(n) && Gen::supported(n)) return Gen::r2c(grid, n, nb);
  dispatch(n, [&](auto impl) { decltype(impl)::r2c(grid, nb); });
}

void fft_c2r_3d(float* grid, int n, int nb) {
  if (!tuned(n) && Gen::supported(n)) return Gen::c2r(grid, n, nb);
  dispatch(n, [&](auto impl) { decltype(impl)::c2r(grid, nb); });
}

void fft_force_meshes(float* mesh, float* const out[3], int n, float f1, int nb) {
  if (!tuned(n) && Gen::supported(n)) return Gen::force_meshes(mesh, out, n, f1, nb);
  dispatch(n, [&](auto impl) { decltype(impl)::force_meshes(mesh, out, f1, nb); });
}

void fft_force_spectra(float* mesh, float* const out[3], int n, float f1, bool z_done) {
  if (!tuned(n)) fail(FS2_ERR_VALUE, "force spectra: %d is not a tuned transform size", n);
  dispatch(n, [&](auto impl) { decltype(impl)::force_spectra(mesh, out, f1, 1, z_done); });
}

void fft_force_finish(float* const out[3], int n) {
  if (!tuned(n)) fail(FS2_ERR_VALUE, "force spectra: %d is not a tuned transform size", n);
  dispatch(n, [&](auto impl) {
    for (int a = 0; a < 3; ++a) decltype(impl)::z_c2r(out[a], 1);
  });
}

const float2* fft_twiddle_table(int N) { return twiddles(N); }

void fft_adjoint_green(float* const p[3], int n, float f1) {
  if (!tuned(n) && Gen::supported(n)) return Gen::adjoint_green(p, n, f1);
  dispatch(n, [&](auto impl) { decltype(impl)::adjoint_green(p, f1); });
}

void fft_green_c2r_3d(const float* delta_k, float* dst, int n, int axis, float f1) {
  if (!tuned(n) && Gen::supported(n)) return Gen::green_c2r(delta_k, dst, n, axis, f1);
  dispatch(n, [&](auto impl) { decltype(impl)::green_c2r(delta_k, dst, axis, f1); });
}

}  // namespace fs2
