// Device-side building blocks of the Stockham FFTs: complex helpers,
// radix-R DFTs, per-length plans, LDS layouts, barrier policies and the
// register-resident transform fft_regs.  fft.hip builds its passes from
// them; paint_tiled.hip uses the z-row c2r in the fused column readout.
#pragma once

#include <hip/hip_runtime.h>

#include <type_traits>

namespace fs2 {
namespace fftdev {

// ---- complex helpers -------------------------------------------------------------
__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
  return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ float2 cconj(float2 a) { return make_float2(a.x, -a.y); }
__device__ __forceinline__ float2 mul_negi(float2 a) { return make_float2(a.y, -a.x); }   // -i a
__device__ __forceinline__ float2 mul_posi(float2 a) { return make_float2(-a.y, a.x); }   //  i a
__device__ __forceinline__ float2 cscale(float2 a, float s) { return make_float2(a.x * s, a.y * s); }

// c2r along z of a row of n = 2M reals: the split of the r2c inverted, with
// the imaginary parts of X[0] and X[M] dropped.  From A = X[k], B = X[M - k]
// and tw = W^k (W = e^{-2 pi i/n}):  Z[k] = E' + i O',  E' = A + conj B,
// O' = (A - conj B) W^{-k}; z = IDFT_M(Z) = x_even + i x_odd.  Returned
// conjugated, since the inverse runs as a conjugated forward transform.
__device__ __forceinline__ float2 c2r_merge(float2 A, float2 B, float2 tw, bool k0) {
  if (k0) {
    A.y = 0.f;
    B.y = 0.f;
  }
  const float2 Bc = cconj(B);
  const float2 E = cadd(A, Bc);
  const float2 O = cmul(csub(A, Bc), cconj(tw));
  return cconj(cadd(E, mul_posi(O)));
}

// r2c along z: a row of n = 2M reals read as M complex z[m] = x[2m] + i
// x[2m+1] and transformed (Z = DFT_M z) is split into the half spectrum:
// X[k] = E - i W^k O,  E = (Z[k] + conj Z[M-k])/2,  O = (Z[k] - conj Z[M-k])/2,
// from A = Z[k], B = Z[M - k] (Z[0] for k = 0) and tw = W^k; X[M] = Re Z[0]
// - Im Z[0].
__device__ __forceinline__ float2 r2c_split(float2 A, float2 B, float2 tw) {
  const float2 Bc = cconj(B);
  const float2 E = cscale(cadd(A, Bc), 0.5f);
  const float2 O = cscale(csub(A, Bc), 0.5f);
  return cadd(E, mul_negi(cmul(tw, O)));
}

// ---- radix-R DFTs, forward sign ----------------------------------------------------
template <int R> struct DFT;

template <> struct DFT<2> {
  static __device__ __forceinline__ void run(float2* v) {
    float2 a = v[0], b = v[1];
    v[0] = cadd(a, b);
    v[1] = csub(a, b);
  }
};

template <> struct DFT<3> {
  static __device__ __forceinline__ void run(float2* v) {
    const float c = -0.5f, s = 0.86602540378443864676f;   // w = c - i s
    float2 a = v[0], sum = cadd(v[1], v[2]), dif = csub(v[1], v[2]);
    float2 m = make_float2(a.x + c * sum.x, a.y + c * sum.y);
    float2 r = make_float2(s * dif.y, -s * dif.x);         // -i s (x1 - x2)
    v[0] = cadd(a, sum);
    v[1] = cadd(m, r);
    v[2] = csub(m, r);
  }
};

template <> struct DFT<4> {
  static __device__ __forceinline__ void run(float2* v) {
    float2 t0 = cadd(v[0], v[2]), t1 = csub(v[0], v[2]);
    float2 t2 = cadd(v[1], v[3]), t3 = mul_negi(csub(v[1], v[3]));
    v[0] = cadd(t0, t2);
    v[2] = csub(t0, t2);
    v[1] = cadd(t1, t3);
    v[3] = csub(t1, t3);
  }
};

template <> struct DFT<8> {
  static __device__ __forceinline__ void run(float2* v) {
    const float h = 0.70710678118654752440f;
    float2 e[4] = {v[0], v[2], v[4], v[6]};
    float2 o[4] = {v[1], v[3], v[5], v[7]};
    DFT<4>::run(e);
    DFT<4>::run(o);
    // w8^k, k = 0..3 : 1, (1-i)/sqrt2, -i, (-1-i)/sqrt2
    float2 o1 = make_float2(h * (o[1].x + o[1].y), h * (o[1].y - o[1].x));
    float2 o2 = mul_negi(o[2]);
    float2 o3 = make_float2(h * (o[3].y - o[3].x), -h * (o[3].x + o[3].y));
    v[0] = cadd(e[0], o[0]); v[4] = csub(e[0], o[0]);
    v[1] = cadd(e[1], o1);   v[5] = csub(e[1], o1);
    v[2] = cadd(e[2], o2);   v[6] = csub(e[2], o2);
    v[3] = cadd(e[3], o3);   v[7] = csub(e[3], o3);
  }
};

template <> struct DFT<16> {
  static __device__ __forceinline__ void run(float2* v) {
    // 4 x 4: A_r = DFT4(x_r, x_{r+4}, x_{r+8}, x_{r+12});
    // y[k1 + 4 k2] = DFT4_r( w16^{r k1} A_r[k1] )[k2]
    float2 a[4][4];
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      a[r][0] = v[r]; a[r][1] = v[r + 4]; a[r][2] = v[r + 8]; a[r][3] = v[r + 12];
      DFT<4>::run(a[r]);
    }
    const float c1 = 0.92387953251128675613f, s1 = 0.38268343236508977173f;
    const float h = 0.70710678118654752440f;
    // w16^m for m = 0..9 (r k1 <= 9): cos(2 pi m/16) - i sin(2 pi m/16)
    const float2 w[10] = {
        make_float2(1.f, 0.f),  make_float2(c1, -s1), make_float2(h, -h),   make_float2(s1, -c1),
        make_float2(0.f, -1.f), make_float2(-s1, -c1), make_float2(-h, -h), make_float2(-c1, -s1),
        make_float2(-1.f, 0.f), make_float2(-c1, s1)};
#pragma unroll
    for (int k1 = 0; k1 < 4; ++k1) {
      float2 b[4];
#pragma unroll
      for (int r = 0; r < 4; ++r) b[r] = (r * k1 == 0) ? a[r][k1] : cmul(a[r][k1], w[r * k1]);
      DFT<4>::run(b);
#pragma unroll
      for (int k2 = 0; k2 < 4; ++k2) v[k1 + 4 * k2] = b[k2];
    }
  }
};

// ---- plans -----------------------------------------------------------------------------
template <int... Rs> struct Radices {};
template <int N> struct Plan;
// T = threads per pencil; every stage radix R must satisfy N % (R T) == 0.
template <> struct Plan<2>    { static constexpr int T = 1;  using R = Radices<2>; };
template <> struct Plan<4>    { static constexpr int T = 1;  using R = Radices<4>; };
template <> struct Plan<8>    { static constexpr int T = 1;  using R = Radices<8>; };
template <> struct Plan<16>   { static constexpr int T = 4;  using R = Radices<4, 4>; };
template <> struct Plan<24>   { static constexpr int T = 2;  using R = Radices<4, 2, 3>; };
template <> struct Plan<32>   { static constexpr int T = 4;  using R = Radices<8, 4>; };
template <> struct Plan<48>   { static constexpr int T = 4;  using R = Radices<4, 4, 3>; };
template <> struct Plan<64>   { static constexpr int T = 8;  using R = Radices<8, 8>; };
template <> struct Plan<96>   { static constexpr int T = 4;  using R = Radices<8, 4, 3>; };
template <> struct Plan<128>  { static constexpr int T = 8;  using R = Radices<16, 8>; };
template <> struct Plan<256>  { static constexpr int T = 16; using R = Radices<16, 16>; };
template <> struct Plan<512>  { static constexpr int T = 64; using R = Radices<8, 8, 8>; };
template <> struct Plan<1024> { static constexpr int T = 64; using R = Radices<16, 16, 4>; };

// LDS layouts of one pencil (element j at s[L::at(j)]).  The padding
// breaks the power-of-two strides of the Stockham writes (rows R apart in
// the first stage) across LDS banks.
//   at(j + m) = at(j) + span(m)  for m a multiple of kPeriod (any j >= 0);
//   at(x + r) = at(x) + r * kUnit  for x a multiple of R, r < R, R | kPeriod;
// the stages use these to address a thread's R elements from one base with
// compile-time offsets (immediate LDS offsets) instead of R index
// computations.
template <int P>
struct RowsLayout {   // [j][P] tiles of strided pencils, one pad row per 16
  static constexpr int kPeriod = 16, kUnit = P;
  static constexpr int rows(int N) { return N + N / 16; }
  static constexpr int span(int m) { return (m + m / 16) * P; }
  static __device__ __forceinline__ int at(int j) { return (j + (j >> 4)) * P; }
};
struct ContigLayout {   // contiguous pencil, one pad element per 8
  static constexpr int kPeriod = 8, kUnit = 1;
  static constexpr int size(int N) { return N + N / 8 + 1; }
  static constexpr int span(int m) { return m + m / 8; }
  static __device__ __forceinline__ int at(int j) { return j + (j >> 3); }
};

// s[at(j + r m)] for r < R, as one base plus constant offsets when the
// layout allows it.
template <class L, int M>
__device__ __forceinline__ int lds_elem(int a0, int j, int r) {
  if constexpr (M % L::kPeriod == 0) {
    return a0 + r * L::span(M);
  } else {
    return L::at(j + r * M);
  }
}

// Barriers between the LDS exchanges of a transform.  Pencils spread over
// several wavefronts need the workgroup barrier.  A row whose threads all sit
// in one wavefront (z passes with T <= 64 threads per row, rows packed 64/T
// to a wavefront, each in its own LDS slice) needs only the wavefront's own
// LDS instructions kept in program order, which the LDS executes in order:
// the waves of a workgroup then run their rows without waiting for each
// other.
struct BlockSync {
  static __device__ __forceinline__ void sync() { __syncthreads(); }
};
struct WaveSync {
  static __device__ __forceinline__ void sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  }
};

// One Stockham stage over one pencil held in LDS at s[L::at(j)].
template <int N, int T, class L, int Ns, int R, class S = BlockSync>
__device__ __forceinline__ void stockham_stage(float2* s, int t, const float2* __restrict__ tw) {
  static_assert(N % (R * T) == 0, "plan: stage does not divide evenly");
  constexpr int B = N / (R * T);
  float2 v[B][R];
#pragma unroll
  for (int b = 0; b < B; ++b) {
    const int j = t + b * T;
    const int a0 = L::at(j);
#pragma unroll
    for (int r = 0; r < R; ++r) v[b][r] = s[lds_elem<L, N / R>(a0, j, r)];
  }
  S::sync();
#pragma unroll
  for (int b = 0; b < B; ++b) {
    const int j = t + b * T;
    const int k = j % Ns;
    if constexpr (Ns > 1) {
#pragma unroll
      for (int r = 1; r < R; ++r) v[b][r] = cmul(v[b][r], tw[r * k * (N / (Ns * R))]);
    }
    DFT<R>::run(v[b]);
    const int base = (j - k) * R + k;
    const int a0 = L::at(base);
#pragma unroll
    for (int r = 0; r < R; ++r) s[lds_elem<L, Ns>(a0, base, r)] = v[b][r];
  }
  S::sync();
}

// Register-resident variant.  Thread t holds elements t + i T (i < N/T) of
// the pencil in x[i]; on return x[i] holds output element t + i T.  The
// first Stockham stage (span 1) reads exactly x[b + r B] (B = N/(R T)) and
// the last stage (span N/R) writes exactly x[b + r B], so only the stores of
// the first stage and the inner stages go through LDS: two LDS passes fewer
// and three barriers instead of 2S + 1.  Starts with a barrier, so the LDS
// pencil may still be in use by the caller on entry.
template <int N, int T, class L, int R, class S = BlockSync>
__device__ __forceinline__ void stage_first(const float2* x, float2* s, int t) {
  constexpr int B = N / (R * T);
#pragma unroll
  for (int b = 0; b < B; ++b) {
    float2 v[R];
#pragma unroll
    for (int r = 0; r < R; ++r) v[r] = x[b + r * B];
    DFT<R>::run(v);
    const int j = t + b * T;
#pragma unroll
    for (int r = 0; r < R; ++r) s[L::at(j * R + r)] = v[r];
  }
  S::sync();
}

template <int N, int T, class L, int Ns, int R>
__device__ __forceinline__ void stage_last(float2* x, const float2* s, int t,
                                           const float2* __restrict__ tw) {
  static_assert(Ns * R == N, "plan: last stage");
  constexpr int B = N / (R * T);
#pragma unroll
  for (int b = 0; b < B; ++b) {
    const int j = t + b * T;   // j < Ns, so the output base is j itself
    float2 v[R];
    const int a0 = L::at(j);
#pragma unroll
    for (int r = 0; r < R; ++r) v[r] = s[lds_elem<L, N / R>(a0, j, r)];
#pragma unroll
    for (int r = 1; r < R; ++r) v[r] = cmul(v[r], tw[r * j * (N / (Ns * R))]);
    DFT<R>::run(v);
#pragma unroll
    for (int r = 0; r < R; ++r) x[b + r * B] = v[r];
  }
}

template <int N, int T, class L, class S, int Ns, int R, int... Rest>
__device__ __forceinline__ void regs_stages(float2* x, float2* s, int t,
                                            const float2* __restrict__ tw) {
  if constexpr (sizeof...(Rest) == 0) {
    stage_last<N, T, L, Ns, R>(x, s, t, tw);
  } else {
    stockham_stage<N, T, L, Ns, R, S>(s, t, tw);
    regs_stages<N, T, L, S, Ns * R, Rest...>(x, s, t, tw);
  }
}

template <int N, int T, class L, class S, int R1, int... Rest>
__device__ __forceinline__ void run_plan_regs(float2* x, float2* s, int t,
                                              const float2* __restrict__ tw, Radices<R1, Rest...>) {
  if constexpr (sizeof...(Rest) == 0) {
    static_assert(T == 1 && R1 == N, "single-stage plans run in one thread");
    DFT<N>::run(x);
  } else {
    S::sync();
    stage_first<N, T, L, R1, S>(x, s, t);
    regs_stages<N, T, L, S, R1, Rest...>(x, s, t, tw);
  }
}

template <int N, class L, class S = BlockSync>
__device__ __forceinline__ void fft_regs(float2* x, float2* s, int t, const float2* __restrict__ tw) {
  run_plan_regs<N, Plan<N>::T, L, S>(x, s, t, tw, typename Plan<N>::R{});
}

// The barrier a z-pass row needs: the wavefront's own when the row's T
// threads sit in one wavefront.
template <int M>
using RowSync = std::conditional_t<(64 % Plan<M>::T == 0), WaveSync, BlockSync>;

// Copies the N-entry twiddle table into LDS (read there by every stage; a
// barrier follows before first use).
template <int N>
__device__ __forceinline__ void stage_twiddles(float2* stw, const float2* __restrict__ tw) {
  for (int i = threadIdx.x; i < N; i += blockDim.x) stw[i] = tw[i];
}

// The same copy in two halves for blocks of THREADS threads: load() before
// the caller's own global loads, store() after them, so that the table's
// latency overlaps theirs (the loop above waits for each table load before
// any later load is issued).  The loads are unconditional, so that no
// branch end waits for them.  A workgroup barrier must follow store()
// before the table is read.
template <int N, int THREADS>
struct TwiddleStage {
  static constexpr int K = (N + THREADS - 1) / THREADS;
  float2 v[K];
  __device__ __forceinline__ void load(const float2* __restrict__ tw) {
#pragma unroll
    for (int k = 0; k < K; ++k) {
      const int i = (int)threadIdx.x + k * THREADS;
      v[k] = tw[i < N ? i : N - 1];
    }
  }
  __device__ __forceinline__ void store(float2* stw) const {
#pragma unroll
    for (int k = 0; k < K; ++k) {
      const int i = (int)threadIdx.x + k * THREADS;
      if (i < N) stw[i] = v[k];
    }
  }
};

}  // namespace fftdev
}  // namespace fs2
