// Tiled CIC mass assignment for gfx950 (replaces the scatter loop of
// py/pm.cpp:57-139 plus the clear of pm.cpp:414-426).
//
// Owner-computes: the mesh is cut into tiles of TX x TY x TZ cells; every
// cell is written once, by the workgroup of the tile that owns it, with a
// plain store of (sum of CIC contributions) - 1.  No global float atomics,
// no separate clear pass.
//
//   K1 bin    : wrap x (pm_domain.cpp:115 fused), lower CIC cell of each
//               particle, key = tile * 8 + spill class; the class bits say
//               whether the lower cell is on the last x / y / z layer of its
//               tile, i.e. whether the +1 corner lands in the next tile.
//               Bin counts are aggregated per workgroup in LDS (one
//               atomic per distinct key and workgroup).
//   K2 scan   : exclusive prefix sum of the counts -> bin offsets.
//   K3 scatter: positions copied into bin order (slots reserved per
//               workgroup: one cursor atomic per distinct key).
//   K4 tile   : per tile, LDS accumulation of every stencil corner that
//               lands in the tile, from the tile's own particles and from the
//               spilling classes of its lower neighbours; then one coalesced
//               store pass.  FP32 LDS atomics (fast variant) or int64
//               fixed-point LDS sums (deterministic variant: integer addition
//               is associative, so the result is bitwise reproducible).
//
// At 1024^3 the tiles are columns of 4 x 2 cells and every z cell: K4 runs
// fused with the density's z r2c (k_tile_paint_r2c), and the readout fused
// with the z c2r of the three force spectra (k_column_readout).  K1 also
// exists with the COLA kick and drift that the previous step deferred in
// front of it (k_kick_drift_bin).
//
// CIC indices and weights follow pm.cpp:85-133 in FP32 with contraction
// disabled, exactly as the per-particle atomic kernel does.
#include <algorithm>
#include <cmath>
#include <type_traits>

#include "fs2_internal.h"
#include "fs2_device.h"
// before the contraction pragma below: the fused column readout's z c2r must
// round exactly as the z c2r pass of fft.hip does
#include "fft_device.h"

#pragma clang fp contract(off)

namespace fs2 {
namespace {

constexpr int kBlock = 256;
// Tile kernels: 512-thread paint and 1024-thread readout blocks keep 32
// waves per CU with their LDS tiles (measured at 1024^3).
constexpr int kTileBlock = 512;
constexpr int kReadoutBlock = 1024;
constexpr unsigned kNoKey = 0xffffffffu;
// log2 of the (x, y) tile-column group edge of the workgroup order
constexpr int kTileGroupLog = 3;

inline unsigned blocks_for(int64_t n) { return (unsigned)((n + kBlock - 1) / kBlock); }

__device__ __forceinline__ float wrap1(float x, float L) {
  if (x < 0.f) x += L;
  if (x >= L) x -= L;
  return x;
}

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

// Tile extents are powers of two for every supported mesh size, so tile
// and cell coordinates come from shifts and masks.
struct TileGeom {
  int n, ncz;
  int tx, ty, tz;      // tile extent in cells
  int lx, ly, lz;      // log2 of the extents
  int ntx, nty, ntz;   // tiles per axis
  float dx_inv;
  // batched launches: nbox boxes of npb particles each (particle i belongs
  // to box i / npb), meshes mstride floats apart, bins_box = 8 * ntiles bins
  // per box; workgroup b works on tile b % ntiles of box b / ntiles.
  int ntiles, nbox;
  unsigned npb, bins_box;
  int64_t mstride;
  // log2 of ntz, nty and ntiles when all three are powers of two, else -1
  int lntz, lnty, lnt;
  // workgroups visit the (x, y) tile columns in 2^lg x 2^lg groups, z
  // fastest, so a tile's +x and +y halo neighbours run close in time on
  // the same XCD and their lines are still in its L2 (0: plain x-y-z order)
  int lg;
};

// Tile of workgroup b (box b / ntiles, tile b % ntiles, z fastest); shifts
// when the tile counts are powers of two, unsigned divisions otherwise.
struct TileCoord {
  int box, tile, Tx, Ty, Tz;
};
__device__ __forceinline__ TileCoord tile_coord(const TileGeom& g, unsigned b) {
  TileCoord c;
  unsigned box, tile, q;
  if (g.lnt >= 0) {
    box = b >> g.lnt;
    const unsigned w = b & ((1u << g.lnt) - 1u);
    c.Tz = (int)(w & ((1u << g.lntz) - 1u));
    q = w >> g.lntz;
    // q = ((Tx_hi * nty_hi + Ty_hi) << 2lg) | (Ty_lo << lg) | Tx_lo
    const unsigned gm = (1u << g.lg) - 1u;
    const unsigned lo = q & ((1u << (2 * g.lg)) - 1u), hi = q >> (2 * g.lg);
    const int lnty_hi = g.lnty - g.lg;
    c.Tx = (int)(((hi >> lnty_hi) << g.lg) | (lo & gm));
    c.Ty = (int)(((hi & ((1u << lnty_hi) - 1u)) << g.lg) | (lo >> g.lg));
    tile = ((unsigned)c.Tx * (unsigned)g.nty + (unsigned)c.Ty) * (unsigned)g.ntz + (unsigned)c.Tz;
  } else {
    box = b / (unsigned)g.ntiles;
    tile = b - box * (unsigned)g.ntiles;
    q = tile / (unsigned)g.ntz;
    c.Tz = (int)(tile - q * (unsigned)g.ntz);
    c.Tx = (int)(q / (unsigned)g.nty);
    c.Ty = (int)(q - (unsigned)c.Tx * (unsigned)g.nty);
  }
  c.box = (int)box;
  c.tile = (int)tile;
  return c;
}

__device__ __forceinline__ unsigned key_of(const TileGeom& g, int ix, int iy, int iz) {
  const unsigned tile = ((unsigned)(ix >> g.lx) * g.nty + (unsigned)(iy >> g.ly)) * g.ntz +
                        (unsigned)(iz >> g.lz);
  const unsigned cls = (((ix & (g.tx - 1)) == g.tx - 1) ? 4u : 0u) |
                       (((iy & (g.ty - 1)) == g.ty - 1) ? 2u : 0u) |
                       (((iz & (g.tz - 1)) == g.tz - 1) ? 1u : 0u);
  return tile * 8u + cls;
}

__global__ void k_fill_u32(unsigned* __restrict__ a, int64_t n, unsigned v) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[i] = v;
}

// Bin of a (wrapped) position: tile of its lower CIC cell and spill class,
// offset by the box's bin range in batched launches.
__device__ __forceinline__ unsigned bin_key(const TileGeom& g, float px, float py, float pz,
                                            int64_t i) {
  const Cic1 cx = cic1(px, g.dx_inv, g.n);
  const Cic1 cy = cic1(py, g.dx_inv, g.n);
  const Cic1 cz = cic1(pz, g.dx_inv, g.n);
  unsigned key = key_of(g, cx.i0, cy.i0, cz.i0);
  if (g.nbox > 1) key += ((unsigned)i / g.npb) * g.bins_box;
  return key;
}

// K1 aggregates its counts per workgroup: kBinPer particles per thread
// (kBinBlock x kBinPer consecutive particles per workgroup).  Within a
// wavefront the lanes form runs of equal keys (consecutive particles are
// neighbouring lattice sites); the first lane of a run finds the key's slot
// in an LDS hash table (linear probing, claimed by compare-and-swap) and
// adds the run's length there, and the workgroup then issues one global
// atomic per distinct key.  The table has twice as many slots as the
// workgroup has particles, so a probe always ends.  (The global atomics,
// not the arithmetic, bound this kernel; the scatter reserves its slots
// over the same table.)
// Clustering breaks up the runs of consecutive particles: at a = 1 a
// workgroup of 1024 particles holds 150 distinct keys against 47 at
// a = 0.1.  Larger workgroups share keys between more of their rows, so
// 4096 particles per workgroup (64 KB of table, two workgroups per CU)
// issue fewer atomics per particle: 10 steps a = 0.1 -> 1 at 512^3 / 1024^3
// bin in 5.5 ms against 6.0 ms with 2048 and 6.9 ms with 1024.
constexpr int kBinBlock = 1024;
constexpr int kBinPer = 4;
constexpr int kBinLogSlots = 13;
constexpr int kBinSlots = 1 << kBinLogSlots;
static_assert(kBinSlots >= 2 * kBinBlock * kBinPer, "bin hash table at most half full");
inline unsigned bin_blocks(int64_t n) {
  return (unsigned)((n + kBinBlock * kBinPer - 1) / (kBinBlock * kBinPer));
}

// On return *start is the lowest lane of this lane's run of equal keys and
// *len the run's length (meaningful on the run's first lane).
__device__ __forceinline__ void key_runs(unsigned key, int* start, int* len) {
  const int lane = __lane_id();
  const unsigned prev = (unsigned)__shfl_up((int)key, 1);
  const uint64_t starts = __ballot(lane == 0 || prev != key);
  // highest run start at or below this lane
  const uint64_t below = starts & (~0ull >> (63 - lane));
  *start = 63 - __builtin_clzll(below);
  // next run start above this lane, or 64
  const uint64_t above = lane == 63 ? 0ull : starts & (~0ull << (lane + 1));
  *len = (above ? __builtin_ctzll(above) : 64) - lane;
}

__device__ __forceinline__ int bin_table_slot(unsigned* tkey, unsigned key) {
  unsigned h = (key * 2654435761u) >> (32 - kBinLogSlots);
  for (;;) {
    const unsigned k = __atomic_load_n(&tkey[h], __ATOMIC_RELAXED);
    if (k == key) return (int)h;
    if (k == kNoKey) {
      const unsigned old = atomicCAS(&tkey[h], kNoKey, key);
      if (old == kNoKey || old == key) return (int)h;
    }
    h = (h + 1u) & (kBinSlots - 1u);
  }
}

// K1: wrap (written back only where it changes x) and bin counts.  The
// scatter recomputes the same key from the written positions, so no key
// array is stored.
__global__ void __launch_bounds__(kBinBlock)
k_bin(float* __restrict__ x, int64_t stride, int64_t np, float L, int do_wrap, TileGeom g,
      unsigned* __restrict__ counts) {
  __shared__ unsigned tkey[kBinSlots], tcnt[kBinSlots];
  for (int e = threadIdx.x; e < kBinSlots; e += blockDim.x) {
    tkey[e] = kNoKey;
    tcnt[e] = 0u;
  }
  __syncthreads();
  const int64_t i0 = (int64_t)blockIdx.x * (kBinBlock * kBinPer) + threadIdx.x;
  // every position load first: the wrap stores below could alias later
  // loads as far as the compiler knows, which would serialise them
  float p[kBinPer][3];
#pragma unroll
  for (int k = 0; k < kBinPer; ++k) {
    const int64_t i = i0 + k * kBinBlock;
#pragma unroll
    for (int a = 0; a < 3; ++a) p[k][a] = i < np ? x[i + a * stride] : 0.f;
  }
#pragma unroll
  for (int k = 0; k < kBinPer; ++k) {
    const int64_t i = i0 + k * kBinBlock;
    unsigned key = kNoKey;
    if (i < np) {
      if (do_wrap) {
#pragma unroll
        for (int a = 0; a < 3; ++a) {
          const float w = wrap1(p[k][a], L);
          if (w != p[k][a]) x[i + a * stride] = w;
          p[k][a] = w;
        }
      }
      key = bin_key(g, p[k][0], p[k][1], p[k][2], i);
    }
    int start, len;
    key_runs(key, &start, &len);
    if (start == __lane_id() && key != kNoKey)
      atomicAdd(&tcnt[bin_table_slot(tkey, key)], (unsigned)len);
  }
  __syncthreads();
  for (int e = threadIdx.x; e < kBinSlots; e += blockDim.x) {
    const unsigned c = tcnt[e];
    if (c) atomicAdd(counts + tkey[e], c);
  }
}

// K1 with the COLA step that precedes it (fs2_cola_drift defers its update
// to the next density binning): each particle's kick, if one is pending, and
// drift as k_kick_drift / k_drift compute them (kick_drift_coord), the
// periodic wrap, one store of the positions, and k_bin's count.  k_bin's
// read of the positions is saved; the velocities are written only when a
// kick ran.
__global__ void __launch_bounds__(kBinBlock)
k_kick_drift_bin(float* __restrict__ x, float* __restrict__ v, const float* __restrict__ f,
                 const float* __restrict__ dx1, const float* __restrict__ dx2, int64_t stride,
                 int64_t np, KickDrift kd, float L, int do_wrap, TileGeom g,
                 unsigned* __restrict__ counts) {
  __shared__ unsigned tkey[kBinSlots], tcnt[kBinSlots];
  for (int e = threadIdx.x; e < kBinSlots; e += blockDim.x) {
    tkey[e] = kNoKey;
    tcnt[e] = 0u;
  }
  __syncthreads();
  const int64_t i0 = (int64_t)blockIdx.x * (kBinBlock * kBinPer) + threadIdx.x;
#pragma unroll
  for (int k = 0; k < kBinPer; ++k) {
    const int64_t i = i0 + k * kBinBlock;
    unsigned key = kNoKey;
    if (i < np) {
      float p[3];
#pragma unroll
      for (int a = 0; a < 3; ++a) {
        const int64_t j = i + a * stride;
        float vj = v[j];
        float xj = kick_drift_coord(kd, x[j], vj, f + j, dx1[j], dx2[j]);
        if (kd.kick) v[j] = vj;
        if (do_wrap) xj = wrap1(xj, L);
        x[j] = xj;
        p[a] = xj;
      }
      key = bin_key(g, p[0], p[1], p[2], i);
    }
    int start, len;
    key_runs(key, &start, &len);
    if (start == __lane_id() && key != kNoKey)
      atomicAdd(&tcnt[bin_table_slot(tkey, key)], (unsigned)len);
  }
  __syncthreads();
  for (int e = threadIdx.x; e < kBinSlots; e += blockDim.x) {
    const unsigned c = tcnt[e];
    if (c) atomicAdd(counts + tkey[e], c);
  }
}

// Exclusive scan of the bin counts in three launches: per-block scans of
// kScanChunk values, a one-block scan of the block totals, and the add-back
// that writes offsets and cursors.
constexpr int kScanThreads = 1024;
constexpr int kScanPer = 4;
constexpr int kScanChunk = kScanThreads * kScanPer;

__device__ __forceinline__ unsigned block_exclusive_scan(unsigned v, unsigned* lds, unsigned* total) {
  const int t = threadIdx.x;
  lds[t] = v;
  __syncthreads();
  for (int off = 1; off < kScanThreads; off <<= 1) {
    const unsigned u = t >= off ? lds[t - off] : 0u;
    __syncthreads();
    lds[t] += u;
    __syncthreads();
  }
  const unsigned incl = lds[t];
  if (total) *total = lds[kScanThreads - 1];
  __syncthreads();
  return incl - v;
}

__global__ void __launch_bounds__(kScanThreads)
k_scan_blocks(const unsigned* __restrict__ counts, int64_t n, unsigned* __restrict__ partial,
              unsigned* __restrict__ block_sums) {
  __shared__ unsigned lds[kScanThreads];
  const int64_t base = (int64_t)blockIdx.x * kScanChunk + (int64_t)threadIdx.x * kScanPer;
  unsigned c[kScanPer];
  unsigned s = 0;
#pragma unroll
  for (int i = 0; i < kScanPer; ++i) {
    c[i] = base + i < n ? counts[base + i] : 0u;
    s += c[i];
  }
  unsigned total;
  unsigned run = block_exclusive_scan(s, lds, &total);
#pragma unroll
  for (int i = 0; i < kScanPer; ++i) {
    if (base + i < n) partial[base + i] = run;
    run += c[i];
  }
  if (threadIdx.x == 0) block_sums[blockIdx.x] = total;
}

// One block: exclusive scan of nb block totals (nb <= kScanChunk), in place;
// offsets[n] receives the grand total.
__global__ void __launch_bounds__(kScanThreads)
k_scan_sums(unsigned* __restrict__ block_sums, int nb, unsigned* __restrict__ offsets, int64_t n) {
  __shared__ unsigned lds[kScanThreads];
  const int base = threadIdx.x * kScanPer;
  unsigned c[kScanPer];
  unsigned s = 0;
#pragma unroll
  for (int i = 0; i < kScanPer; ++i) {
    c[i] = base + i < nb ? block_sums[base + i] : 0u;
    s += c[i];
  }
  unsigned total;
  unsigned run = block_exclusive_scan(s, lds, &total);
#pragma unroll
  for (int i = 0; i < kScanPer; ++i) {
    if (base + i < nb) block_sums[base + i] = run;
    run += c[i];
  }
  if (threadIdx.x == 0) offsets[n] = total;
}

// partial and cursor may be the same buffer (same index read then written).
__global__ void k_scan_add(const unsigned* partial, const unsigned* __restrict__ block_sums,
                           int64_t n, unsigned* __restrict__ offsets, unsigned* cursor) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const unsigned v = partial[i] + block_sums[i / kScanChunk];
  offsets[i] = v;
  cursor[i] = v;
}

// One 16-byte record per particle in bin order: position and the particle's
// index (bits in .w), which the tiled readout uses to write forces back in
// id order.  Slots are reserved per workgroup, over the particles and the
// table of K1: a run of equal keys in a wavefront takes its ranks within the
// workgroup from one LDS atomic, the workgroup reserves one contiguous range
// per distinct key with one global atomic, and each particle stores its
// record at the range start plus its rank.
__global__ void __launch_bounds__(kBinBlock)
k_scatter(const float* __restrict__ x, int64_t stride, int64_t np, TileGeom g,
          unsigned* __restrict__ cursor, float4* __restrict__ rec) {
  __shared__ unsigned tkey[kBinSlots], tcnt[kBinSlots];
  for (int e = threadIdx.x; e < kBinSlots; e += blockDim.x) {
    tkey[e] = kNoKey;
    tcnt[e] = 0u;
  }
  __syncthreads();
  const int64_t i0 = (int64_t)blockIdx.x * (kBinBlock * kBinPer) + threadIdx.x;
  float p[kBinPer][3];
#pragma unroll
  for (int k = 0; k < kBinPer; ++k) {
    const int64_t i = i0 + k * kBinBlock;
#pragma unroll
    for (int a = 0; a < 3; ++a) p[k][a] = i < np ? x[i + a * stride] : 0.f;
  }
  const int lane = __lane_id();
  int slot[kBinPer];
  unsigned rank[kBinPer];
#pragma unroll
  for (int k = 0; k < kBinPer; ++k) {
    const int64_t i = i0 + k * kBinBlock;
    const unsigned key = i < np ? bin_key(g, p[k][0], p[k][1], p[k][2], i) : kNoKey;
    int start, len;
    key_runs(key, &start, &len);
    int sl = -1;
    unsigned r0 = 0u;
    if (start == lane && key != kNoKey) {
      sl = bin_table_slot(tkey, key);
      r0 = atomicAdd(&tcnt[sl], (unsigned)len);
    }
    sl = __shfl(sl, start);
    r0 = (unsigned)__shfl((int)r0, start);
    slot[k] = key != kNoKey ? sl : -1;
    rank[k] = r0 + (unsigned)(lane - start);
  }
  __syncthreads();
  // one global atomic per distinct key; the count becomes the range start
  for (int e = threadIdx.x; e < kBinSlots; e += blockDim.x) {
    const unsigned c = tcnt[e];
    if (c) tcnt[e] = atomicAdd(cursor + tkey[e], c);
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < kBinPer; ++k) {
    if (slot[k] >= 0) {
      const int64_t i = i0 + k * kBinBlock;
      // the index fits in 32 bits: 2^32 particles (68 bytes each) exceed the HBM
      rec[tcnt[slot[k]] + rank[k]] =
          make_float4(p[k][0], p[k][1], p[k][2], __uint_as_float((unsigned)i));
    }
  }
}

// Source tiles S = T - d, d in {0,1}^3 (periodic).  A particle of S
// reaches T only by spilling along every axis where S != T, so S
// contributes the classes c with (c & d) == d.  An offset along an axis with
// a single tile names T's own column again; such d are skipped, their
// particles being those of the d without that bit.  The bins of one tile
// are contiguous, so a run of consecutive selected classes is one record
// range: at most 14 ranges, the first (d = 0) holding all of T's own
// records.  Wavefront 0 examines the 64 (d, class) pairs, one per lane, and
// compacts the run starts: range r covers records [rb[r], rb[r] + rpre[r+1]
// - rpre[r]), rpre is their exclusive prefix sum and rpre[nranges] the
// total.  rb needs 64 entries, rpre 65.  Block-cooperative; ends with a
// barrier.

// The two bin offsets that bound lane (d, cl)'s range (loads only, so that
// a caller can issue other loads before it uses them; act = false: no run
// of selected classes starts at this lane).
struct RangeBounds {
  unsigned bb = 0, e = 0;
  bool act = false;
};
__device__ __forceinline__ RangeBounds source_bounds_lane(const unsigned* __restrict__ offsets,
                                                          const TileGeom& g, int Tx, int Ty,
                                                          int Tz, int lane) {
  const int d = lane >> 3, cl = lane & 7;
  const int single = (g.ntx == 1 ? 4 : 0) | (g.nty == 1 ? 2 : 0) | (g.ntz == 1 ? 1 : 0);
  auto sel = [d](int c) { return (c & d) == d; };
  int end = cl;
  while (end < 7 && sel(end + 1)) ++end;
  const unsigned Sx = (d & 4) ? (Tx ? Tx - 1 : g.ntx - 1) : Tx;
  const unsigned Sy = (d & 2) ? (Ty ? Ty - 1 : g.nty - 1) : Ty;
  const unsigned Sz = (d & 1) ? (Tz ? Tz - 1 : g.ntz - 1) : Tz;
  const unsigned S8 = ((Sx * (unsigned)g.nty + Sy) * (unsigned)g.ntz + Sz) * 8u;
  // every lane loads (S8 + end + 1 <= 8 ntiles, the last offset), so that
  // the loads sit behind no branch whose end would have to wait for them
  RangeBounds r;
  r.bb = offsets[S8 + cl];
  r.e = offsets[S8 + end + 1];
  r.act = (d & single) == 0 && sel(cl) && (cl == 0 || !sel(cl - 1));
  return r;
}
__device__ __forceinline__ void source_range_of(const RangeBounds& r, unsigned* b,
                                                unsigned* len) {
  *b = r.act ? r.bb : 0u;
  *len = r.act && r.e > r.bb ? r.e - r.bb : 0u;
}
// Range of lane (d, cl) = (lane / 8, lane % 8) of tile_source_ranges: the
// record range [b, b + len) of source tile T - d starting at class cl when a
// run of selected classes starts there, else len = 0.
__device__ __forceinline__ void source_range_lane(const unsigned* __restrict__ offsets,
                                                  const TileGeom& g, int Tx, int Ty, int Tz,
                                                  int lane, unsigned* b, unsigned* len) {
  source_range_of(source_bounds_lane(offsets, g, Tx, Ty, Tz, lane), b, len);
}

__device__ __forceinline__ void tile_source_ranges(const unsigned* __restrict__ offsets,
                                                   const TileGeom& g, int Tx, int Ty, int Tz,
                                                   unsigned* rb, unsigned* rpre, int* nranges) {
  if (threadIdx.x < 64) {
    const int lane = threadIdx.x;
    unsigned b, len;
    source_range_lane(offsets, g, Tx, Ty, Tz, lane, &b, &len);
    unsigned incl = len;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
      const unsigned y = __shfl_up(incl, off);
      if (lane >= off) incl += y;
    }
    const uint64_t nonempty = __ballot(len > 0);
    const int k = __popcll(nonempty & ((1ull << lane) - 1ull));
    if (len > 0) {
      rb[k] = b;
      rpre[k] = incl - len;
    }
    if (lane == 63) {
      *nranges = __popcll(nonempty);
      rpre[__popcll(nonempty)] = incl;
    }
  }
  __syncthreads();
}

// Zeroes the first `bytes` bytes of LDS (a multiple of 4) with 16-byte
// stores and a word-wise tail.  The caller's next barrier orders it.
__device__ __forceinline__ void lds_zero(unsigned char* smem, int bytes) {
  uint4* a4 = (uint4*)smem;
  const int n16 = bytes >> 4;
  for (int i = threadIdx.x; i < n16; i += blockDim.x) a4[i] = make_uint4(0u, 0u, 0u, 0u);
  const int tail = (bytes & 15) >> 2;
  if ((int)threadIdx.x < tail) ((unsigned*)(smem + (n16 << 4)))[threadIdx.x] = 0u;
}

// Offset of owned cell c (z fastest within the tile) from the tile's origin
// cell in a padded mesh: (lx * n + ly) * ncz + lz, below 2^31 for any tile.
__device__ __forceinline__ int tile_cell_offset(const TileGeom& g, int c) {
  const int lz = c & (g.tz - 1);
  const int row = c >> g.lz;
  return ((row >> g.ly) * g.n + (row & (g.ty - 1))) * g.ncz + lz;
}

// Visits the owned cells of a tile in groups of V = 4, 2 or 1 consecutive z
// cells, fn(c, offset, V) with c the first cell.  Rows of the padded mesh
// start at multiples of 32 floats and z0 is a multiple of tz, so a group
// of V cells is one aligned V-float store when V divides tz.
template <typename Fn>
__device__ __forceinline__ void for_tile_groups(const TileGeom& g, Fn fn) {
  const int tcells = g.tx * g.ty * g.tz;
  if ((g.tz & 3) == 0) {
    for (int c = 4 * threadIdx.x; c < tcells; c += 4 * blockDim.x)
      fn(c, tile_cell_offset(g, c), std::integral_constant<int, 4>());
  } else if ((g.tz & 1) == 0) {
    for (int c = 2 * threadIdx.x; c < tcells; c += 2 * blockDim.x)
      fn(c, tile_cell_offset(g, c), std::integral_constant<int, 2>());
  } else {
    for (int c = threadIdx.x; c < tcells; c += blockDim.x)
      fn(c, tile_cell_offset(g, c), std::integral_constant<int, 1>());
  }
}

// Stores V consecutive floats v[0..V) at p (V-float aligned).
template <int V>
__device__ __forceinline__ void store_group(float* p, const float* v) {
  if constexpr (V == 4) {
    *(float4*)p = make_float4(v[0], v[1], v[2], v[3]);
  } else if constexpr (V == 2) {
    *(float2*)p = make_float2(v[0], v[1]);
  } else {
    *p = v[0];
  }
}

// Accumulates every CIC corner that lands in tile tc into acc (LDS, zeroed by
// the caller), from the tile's own particles and from the spilling classes
// of its lower neighbours (tile_source_ranges; rb, rpre, nranges are its
// shared arrays).  ACC is float (LDS float atomics) or unsigned long long
// (int64 fixed point, deterministic).  Ends with a barrier.
// Adds the CIC corners of one record that land in the tile with origin
// cell (x0, y0, z0) to acc (f0 = fac, times the fixed-point scale for the
// int64 accumulator).
template <typename ACC>
__device__ __forceinline__ void deposit_record(ACC* acc, const float4 r4, const TileGeom& g,
                                               int x0, int y0, int z0, float f0) {
  const Cic1 cx = cic1(r4.x, g.dx_inv, g.n);
  const Cic1 cy = cic1(r4.y, g.dx_inv, g.n);
  const Cic1 cz = cic1(r4.z, g.dx_inv, g.n);
  const int lx[2] = {cx.i0 - x0, cx.i1 - x0};
  const int ly[2] = {cy.i0 - y0, cy.i1 - y0};
  const int lz[2] = {cz.i0 - z0, cz.i1 - z0};
  const float wx[2] = {cx.w0, cx.w1};
  const float wy[2] = {cy.w0, cy.w1};
  const float wz[2] = {cz.w0, cz.w1};
#pragma unroll
  for (int ex = 0; ex < 2; ++ex) {
    if ((unsigned)lx[ex] >= (unsigned)g.tx) continue;
#pragma unroll
    for (int ey = 0; ey < 2; ++ey) {
      if ((unsigned)ly[ey] >= (unsigned)g.ty) continue;
#pragma unroll
      for (int ez = 0; ez < 2; ++ez) {
        if ((unsigned)lz[ez] >= (unsigned)g.tz) continue;
        // fac*wx*wy*wz, left to right (pm.cpp:122-132)
        const float w = f0 * wx[ex] * wy[ey] * wz[ez];
        const int li = (((lx[ex] << g.ly) + ly[ey]) << g.lz) + lz[ez];
        if constexpr (sizeof(ACC) == 4) {
          atomicAdd(&acc[li], w);
        } else {
          // w <= fac fix <= 2^30: one rounding to int32 per term
          atomicAdd(&acc[li], (ACC)(long long)__float2int_rn(w));
        }
      }
    }
  }
}

template <typename ACC>
__device__ __forceinline__ void tile_deposit(ACC* acc, const float4* __restrict__ rec,
                                             const unsigned* __restrict__ offsets,
                                             const TileGeom& g, const TileCoord& tc, float fac,
                                             float fix, unsigned* rb, unsigned* rpre,
                                             int* nranges) {
  const int x0 = tc.Tx * g.tx, y0 = tc.Ty * g.ty, z0 = tc.Tz * g.tz;
  tile_source_ranges(offsets, g, tc.Tx, tc.Ty, tc.Tz, rb, rpre, nranges);
  const int nr = *nranges;
  const unsigned total = rpre[nr];
  // fixed point: fix is a power of two, so (fac fix) wx wy wz rounds exactly
  // as (fac wx wy wz) fix
  const float f0 = sizeof(ACC) == 4 ? fac : fac * fix;
  auto deposit = [&](const float4 r4) { deposit_record<ACC>(acc, r4, g, x0, y0, z0, f0); };
  // two records per thread per round, both loads issued before either is
  // deposited (a tile holds ~1.3 records per thread at 1 particle per 8
  // cells, so most threads need one round)
  int r = 0;
  for (unsigned v0 = threadIdx.x; v0 < total; v0 += 2 * blockDim.x) {
    float4 q[2];
    bool ok[2];
#pragma unroll
    for (int u = 0; u < 2; ++u) {
      const unsigned v = v0 + u * blockDim.x;
      ok[u] = v < total;
      if (ok[u]) {
        while (v >= rpre[r + 1]) ++r;
        q[u] = rec[rb[r] + (v - rpre[r])];
      }
    }
#pragma unroll
    for (int u = 0; u < 2; ++u)
      if (ok[u]) deposit(q[u]);
  }
  __syncthreads();
}

// The value a tile cell stores: its CIC sum minus 1 (pm.cpp:414-426 clears to
// -1 before the adds).
template <typename ACC>
__device__ __forceinline__ float cell_value(ACC a, double fix_inv) {
  if constexpr (sizeof(ACC) == 4) {
    return a - 1.0f;
  } else {
    return (float)((double)(long long)a * fix_inv) - 1.0f;
  }
}

// K4: one workgroup per tile.
template <typename ACC>
__global__ void __launch_bounds__(kTileBlock)
k_tile_paint(const float4* __restrict__ rec, const unsigned* __restrict__ offsets,
             TileGeom g, float fac, float fix, double fix_inv, float* __restrict__ grid) {
  extern __shared__ unsigned char smem[];
  ACC* acc = (ACC*)smem;
  const int tcells = g.tx * g.ty * g.tz;
  lds_zero(smem, tcells * (int)sizeof(ACC));

  // z-adjacent tiles (consecutive indices) on one XCD share halo lines in L2
  const TileCoord tc = tile_coord(g, xcd_remap(blockIdx.x, gridDim.x));
  offsets += (int64_t)tc.box * g.bins_box;
  const int x0 = tc.Tx * g.tx, y0 = tc.Ty * g.ty, z0 = tc.Tz * g.tz;
  float* gt = grid + (int64_t)tc.box * g.mstride + ((int64_t)x0 * g.n + y0) * g.ncz + z0;

  __shared__ unsigned rb[64], rpre[65];
  __shared__ int nranges;
  tile_deposit<ACC>(acc, rec, offsets, g, tc, fac, fix, rb, rpre, &nranges);
  for_tile_groups(g, [&](int c, int off, auto vc) {
    constexpr int V = decltype(vc)::value;
    float o[V];
#pragma unroll
    for (int k = 0; k < V; ++k) o[k] = cell_value<ACC>(acc[c + k], fix_inv);
    store_group<V>(gt + off, o);
  });
}

// K4 fused with the z r2c of the force transform (1024^3 meshes, tiles of
// at most kR2cRowWaves full z rows): after the deterministic deposit,
// wavefront w < rows takes row w of the tile, forms its cell values exactly
// as k_tile_paint stores them, transforms the row (its own int64
// accumulators serve as the transform's LDS scratch) and writes the half
// spectrum, as the z r2c pass would from the stored mesh.  The real density
// is never written.
//
// A tile's deposit waits on two dependent loads (its bin offsets, then its
// records) and the 64 KB accumulator of an 8-row tile leaves room for only
// two workgroups per CU, so each workgroup walks kR2cTiles tiles (strided
// by the grid, so that the tiles in flight together stay neighbours: 2.6 GB
// read per launch against 3.8 GB for consecutive tiles, same time) and
// pipelines them: while wavefronts 0-7 transform and write tile
// k (then clear their rows), wavefronts 8-15 load up to kR2cPre records
// each of tile k + 1 into registers, whose source ranges they found one
// tile earlier, and the bin offsets of tile k + 2, issued first, so that
// the two loads of the chain overlap.  The next deposit starts with its
// records in hand; records beyond the prefetch are read from the range
// table wavefront 8 leaves in LDS (one table per tile in flight).  At
// 512^3 particles on the clustered run: 2.90 ms per launch unpipelined;
// 2.53, 2.42-2.44, 2.37 and 2.43 ms with 4, 8, 16 and 32 tiles per
// workgroup and tile k + 1's offsets and records loaded in turn (3 records
// prefetched per thread; 2 records: 2.46-2.48 at 8 tiles); 2.31-2.36 ms
// against 2.40-2.44 with the offsets one tile further ahead.
constexpr int kTileR2cBlock = 1024;
constexpr int kR2cRowWaves = 8;
constexpr unsigned kR2cLoaders = kTileR2cBlock - kR2cRowWaves * 64;
constexpr int kR2cPre = 3;
constexpr int kR2cTiles = 16;

// A tile's source ranges as every loader wavefront holds them: this lane's
// range start b and exclusive prefix, the lanes with a range, the total.
struct R2cRanges {
  unsigned b = 0, excl = 0, total = 0;
  uint64_t nonempty = 0;
};

// The wavefront scan of the lanes' ranges (b, len); `table` wavefronts also
// write them to rb / rpre / nranges as tile_source_ranges does.
__device__ __forceinline__ R2cRanges r2c_scan(unsigned b, unsigned len, int lane, bool table,
                                              unsigned* rb, unsigned* rpre, int* nranges) {
  unsigned incl = len;
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const unsigned y = __shfl_up(incl, off);
    if (lane >= off) incl += y;
  }
  R2cRanges rs;
  rs.b = b;
  rs.excl = incl - len;
  rs.nonempty = __ballot(len > 0);
  rs.total = (unsigned)__builtin_amdgcn_readfirstlane(__shfl((int)incl, 63));
  if (table) {
    const int k = __popcll(rs.nonempty & ((1ull << lane) - 1ull));
    if (len > 0) {
      rb[k] = b;
      rpre[k] = rs.excl;
    }
    if (lane == 63) {
      *nranges = __popcll(rs.nonempty);
      rpre[__popcll(rs.nonempty)] = incl;
    }
  }
  return rs;
}

// Issues the loads of loader thread lt's first kR2cPre records of a tile
// (index lt + i kR2cLoaders; a record's range is found with scalar reads
// of the range starts).
__device__ __forceinline__ void r2c_records(const float4* __restrict__ rec, const R2cRanges& rs,
                                            unsigned lt, float4* pre) {
  unsigned src[kR2cPre];
#pragma unroll
  for (int i = 0; i < kR2cPre; ++i) src[i] = 0u;
  // the ranges in order: record r lies in the last range whose start <= r
  for (uint64_t m = rs.nonempty; m; m &= m - 1ull) {
    const int sl = __builtin_ctzll(m);
    const unsigned p0 = (unsigned)__builtin_amdgcn_readlane((int)rs.excl, sl);
    const unsigned b0 = (unsigned)__builtin_amdgcn_readlane((int)rs.b, sl);
#pragma unroll
    for (int i = 0; i < kR2cPre; ++i) {
      const unsigned r = lt + (unsigned)i * kR2cLoaders;
      if (r >= p0) src[i] = b0 + (r - p0);
    }
  }
  // every thread loads: one past the tile's records reads record 0 (the
  // record scratch holds at least one) and does not deposit it.  No load
  // is behind a branch, so the caller's next wait can leave them in flight.
#pragma unroll
  for (int i = 0; i < kR2cPre; ++i) {
    const unsigned r = lt + (unsigned)i * kR2cLoaders;
    pre[i] = rec[r < rs.total ? src[i] : 0u];
  }
}

template <int M>
__global__ void __launch_bounds__(kTileR2cBlock) __attribute__((amdgpu_waves_per_eu(8)))
k_tile_paint_r2c(const float4* __restrict__ rec, const unsigned* __restrict__ offsets,
                 TileGeom g, float fac, float fix, double fix_inv,
                 const float2* __restrict__ twM, const float2* __restrict__ twN,
                 float2* __restrict__ grid) {
  using namespace fftdev;
  using ACC = unsigned long long;
  constexpr int T = Plan<M>::T;
  static_assert(T == 64, "one wavefront per z row");
  using L = ContigLayout;
  static_assert(L::size(M) * sizeof(float2) <= 2 * M * sizeof(ACC), "row scratch fits a row");
  extern __shared__ unsigned char smem[];
  ACC* acc = (ACC*)smem;
  __shared__ float2 stM[M];
  __shared__ float2 stN[M];
  // range tables of tiles k and k + 1 (tile k's in rb[k & 1])
  __shared__ unsigned rb[2][64], rpre[2][65];
  __shared__ int nranges[2];
  for (int i = threadIdx.x; i < M; i += blockDim.x) {
    stM[i] = twM[i];
    stN[i] = twN[i];
  }
  const int tcells = g.tx * g.ty * g.tz;   // tz = n = 2M: tx * ty rows
  const int rows = g.tx * g.ty;
  lds_zero(smem, tcells * (int)sizeof(ACC));
  const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
  const bool loader = w >= kR2cRowWaves;
  const bool table = w == kR2cRowWaves;
  const unsigned lt = threadIdx.x - (unsigned)(kR2cRowWaves * 64);
  // tile k of this workgroup: k * gridDim.x + its (XCD-remapped) index, so
  // that the tiles in flight at one time are neighbours, as in a one-shot
  // launch, and find the spilling records of their lower neighbours in L2
  const unsigned vb = (unsigned)xcd_remap(blockIdx.x, gridDim.x), vs = gridDim.x;
  // fixed point: fix is a power of two, so (fac fix) wx wy wz rounds exactly
  // as (fac wx wy wz) fix
  const float f0 = fac * fix;
  float4 pre[kR2cPre];
#pragma unroll
  for (int i = 0; i < kR2cPre; ++i) pre[i] = make_float4(0.f, 0.f, 0.f, 0.f);
  R2cRanges cur, nxt;   // the ranges of tiles k and k + 1
  if (loader) {
    unsigned b, len;
    const TileCoord t0 = tile_coord(g, vb);
    source_range_lane(offsets, g, t0.Tx, t0.Ty, t0.Tz, lane, &b, &len);
    cur = r2c_scan(b, len, lane, table, rb[0], rpre[0], &nranges[0]);
    r2c_records(rec, cur, lt, pre);
    if (kR2cTiles > 1) {
      const TileCoord t1 = tile_coord(g, vb + vs);
      source_range_lane(offsets, g, t1.Tx, t1.Ty, t1.Tz, lane, &b, &len);
      nxt = r2c_scan(b, len, lane, table, rb[1], rpre[1], &nranges[1]);
    }
  }
  __syncthreads();   // accumulators clear, twiddles and the first two tables staged
#pragma unroll 1
  for (int k = 0; k < kR2cTiles; ++k) {
    const int tb = k & 1;
    const TileCoord tc = tile_coord(g, vb + (unsigned)k * vs);
    const int x0 = tc.Tx * g.tx, y0 = tc.Ty * g.ty, z0 = tc.Tz * g.tz;
    if (loader) {
#pragma unroll
      for (int i = 0; i < kR2cPre; ++i)
        if (lt + (unsigned)i * kR2cLoaders < cur.total)
          deposit_record<ACC>(acc, pre[i], g, x0, y0, z0, f0);
    }
    // records beyond the prefetch, located through the LDS range table
    const unsigned tot = rpre[tb][nranges[tb]];
    int r = 0;
    for (unsigned v = kR2cLoaders * kR2cPre + threadIdx.x; v < tot; v += blockDim.x) {
      while (v >= rpre[tb][r + 1]) ++r;
      deposit_record<ACC>(acc, rec[rb[tb][r] + (v - rpre[tb][r])], g, x0, y0, z0, f0);
    }
    __syncthreads();   // the tile's cells are complete
    if (!loader) {
      if (w < rows) {
        // row w of the tile: cells (x0 + w / ty, y0 + w % ty, all z)
        const int lx = w >> g.ly, ly = w & (g.ty - 1);
        ACC* ar = acc + (w << g.lz);
        // an opaque copy of the lane: without it the compiler keeps every
        // LDS address of the transform live across the tile loop
        int t = threadIdx.x & 63;
        asm volatile("" : "+v"(t));
        float2 z[M / T];
#pragma unroll
        for (int i = 0; i < M / T; ++i) {
          const int m = t + i * T;
          z[i] = make_float2(cell_value<ACC>(ar[2 * m], fix_inv),
                             cell_value<ACC>(ar[2 * m + 1], fix_inv));
        }
        float2* sp = (float2*)ar;
        fft_regs<M, L, WaveSync>(z, sp, t, stM);   // starts with a wavefront barrier
        WaveSync::sync();   // the row's lanes are done reading the last stage's input
#pragma unroll
        for (int i = 0; i < M / T; ++i) sp[L::at(t + i * T)] = z[i];
        WaveSync::sync();
        float2* out = grid + ((int64_t)(x0 + lx) * g.n + (y0 + ly)) * (g.ncz / 2);
#pragma unroll
        for (int i = 0; i < M / T; ++i) {
          const int kk = t + i * T;
          out[kk] = r2c_split(sp[L::at(kk)], sp[L::at(kk == 0 ? 0 : M - kk)], stN[kk]);
        }
        if (t == 0) {
          const float2 A = sp[L::at(0)];
          out[M] = make_float2(A.x - A.y, 0.f);
        }
        WaveSync::sync();   // the row is read before it is cleared for the next tile
        uint4* a4 = (uint4*)ar;
#pragma unroll
        for (int i = 0; i < 2 * M * (int)sizeof(ACC) / 16 / 64; ++i)
          a4[t + i * 64] = make_uint4(0u, 0u, 0u, 0u);
      }
    } else if (k + 1 < kR2cTiles) {
      // tile k + 2's bin offsets are loaded first, then tile k + 1's records
      // (its ranges are known), then the offsets are used: the next deposit
      // waits on one load latency, not on two in a row.  Tile k + 2's table
      // replaces tile k's, which this tile's deposit finished with above.
      const bool two = k + 2 < kR2cTiles;
      const TileCoord t2 = tile_coord(g, vb + (unsigned)(two ? k + 2 : k + 1) * vs);
      const RangeBounds rk2 = source_bounds_lane(offsets, g, t2.Tx, t2.Ty, t2.Tz, lane);
      r2c_records(rec, nxt, lt, pre);
      cur = nxt;
      if (two) {
        unsigned b, len;
        source_range_of(rk2, &b, &len);
        nxt = r2c_scan(b, len, lane, table, rb[tb], rpre[tb], &nranges[tb]);
      }
    }
    __syncthreads();   // rows written and cleared, tile k + 2's table staged
  }
}

// Copies the (tx+1) x (ty+1) x (tz+1) halo block of NM meshes starting at
// cell (x0, y0, z0) (periodic) into LDS, hs[a * hcells + r * hz + lz] with
// row r = lx * hy + ly.  Mesh rows are tabulated once in LDS (rowoff, one
// int row index per row; offsets in 64 bits), so the main copy of the tz interior columns is division
// free (tz is a power of two) and coalesced along z; the one halo column
// lz = tz is copied by a second, per-row pass.  Ends with a barrier.
template <int NM>
__device__ __forceinline__ void stage_halo(float* hs, int* rowoff, const float* const* ms,
                                           const TileGeom& g, int x0, int y0, int z0) {
  const int hy = g.ty + 1, hz = g.tz + 1;
  const int rows = (g.tx + 1) * hy;
  const int hcells = rows * hz;
  for (int r = threadIdx.x; r < rows; r += blockDim.x) {
    const int lx = r / hy, ly = r - lx * hy;
    int gx = x0 + lx, gy = y0 + ly;
    if (gx >= g.n) gx -= g.n;
    if (gy >= g.n) gy -= g.n;
    rowoff[r] = gx * g.n + gy;   // row index; offsets are formed in 64 bits
  }
  __syncthreads();
  // U elements per thread per round, all loads issued before any LDS store,
  // so a 1024-thread block keeps ~U * NM loads per thread in flight
  constexpr int U = 6;
  const int interior = rows << g.lz;
  for (int i0 = threadIdx.x; i0 < interior; i0 += U * blockDim.x) {
    float v[U][NM];
    int dst[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int i = i0 + u * blockDim.x;
      dst[u] = -1;
      if (i < interior) {
        const int r = i >> g.lz, lz = i & (g.tz - 1);
        const int64_t src = (int64_t)rowoff[r] * g.ncz + z0 + lz;
#pragma unroll
        for (int a = 0; a < NM; ++a) v[u][a] = ms[a][src];
        dst[u] = r * hz + lz;
      }
    }
#pragma unroll
    for (int u = 0; u < U; ++u)
      if (dst[u] >= 0) {
#pragma unroll
        for (int a = 0; a < NM; ++a) hs[a * hcells + dst[u]] = v[u][a];
      }
  }
  int gz = z0 + g.tz;
  if (gz >= g.n) gz -= g.n;
  for (int r = threadIdx.x; r < rows; r += blockDim.x) {
    const int64_t src = (int64_t)rowoff[r] * g.ncz + gz;
#pragma unroll
    for (int a = 0; a < NM; ++a) hs[a * hcells + r * hz + g.tz] = ms[a][src];
  }
  __syncthreads();
}

// LDS bytes of a halo tile of nm meshes plus its row table.
inline size_t halo_lds(const TileGeom& g, int nm) {
  const size_t rows = (size_t)(g.tx + 1) * (g.ty + 1);
  return (size_t)nm * rows * (g.tz + 1) * sizeof(float) + rows * sizeof(int);
}

// Tiled CIC readout (force_at_particle_locations, py/pm.cpp:142-220, all
// three axes): one workgroup per tile stages the three force meshes over the
// tile plus its +1 halo in LDS, then interpolates at the tile's own
// particles (every particle is in exactly one tile's bins) and writes the
// forces back to the particles' id-order slots.  Same FP32 expression order
// as the gather kernel.
__global__ void __launch_bounds__(kReadoutBlock)
k_tile_readout(const float4* __restrict__ rec, const unsigned* __restrict__ offsets, TileGeom g,
               const float* __restrict__ m0, const float* __restrict__ m1,
               const float* __restrict__ m2, float* __restrict__ f, int64_t fstride) {
  extern __shared__ float hs[];
  const int hx = g.tx + 1, hy = g.ty + 1, hz = g.tz + 1;
  const int hcells = hx * hy * hz;
  // z-adjacent tiles (consecutive indices) on one XCD share halo lines in L2
  const TileCoord tc = tile_coord(g, xcd_remap(blockIdx.x, gridDim.x));
  const unsigned tile = (unsigned)tc.tile;
  offsets += (int64_t)tc.box * g.bins_box;
  const int64_t moff = (int64_t)tc.box * g.mstride;
  const int Tx = tc.Tx, Ty = tc.Ty, Tz = tc.Tz;
  const int x0 = Tx * g.tx, y0 = Ty * g.ty, z0 = Tz * g.tz;
  const float* ms[3] = {m0 + moff, m1 + moff, m2 + moff};
  stage_halo<3>(hs, (int*)(hs + 3 * hcells), ms, g, x0, y0, z0);
  const unsigned b = offsets[(unsigned)tile * 8u], e = offsets[(unsigned)tile * 8u + 8u];
  for (unsigned j = b + threadIdx.x; j < e; j += blockDim.x) {
    const float4 r4 = rec[j];
    const Cic1 cx = cic1(r4.x, g.dx_inv, g.n);
    const Cic1 cy = cic1(r4.y, g.dx_inv, g.n);
    const Cic1 cz = cic1(r4.z, g.dx_inv, g.n);
    const int lx = cx.i0 - x0, ly = cy.i0 - y0, lz = cz.i0 - z0;
    const int c000 = (lx * hy + ly) * hz + lz;
    const int c010 = c000 + hz, c100 = c000 + hy * hz, c110 = c100 + hz;
    const unsigned i = __float_as_uint(r4.w);
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      const float* h = hs + a * hcells;
      // pm.cpp:203-218: f = 0; f += (ix0 terms); f += (ix1 terms)
      float s = 0.f;
      s += h[c000] * cx.w0 * cy.w0 * cz.w0 + h[c000 + 1] * cx.w0 * cy.w0 * cz.w1 +
           h[c010] * cx.w0 * cy.w1 * cz.w0 + h[c010 + 1] * cx.w0 * cy.w1 * cz.w1;
      s += h[c100] * cx.w1 * cy.w0 * cz.w0 + h[c100 + 1] * cx.w1 * cy.w0 * cz.w1 +
           h[c110] * cx.w1 * cy.w1 * cz.w0 + h[c110 + 1] * cx.w1 * cy.w1 * cz.w1;
      f[i + a * fstride] = s;
    }
  }
}

// ---- fused z c2r + readout (columns of the 1024^3 mesh) -------------------------------
// The last three passes of the force transform (z c2r of F_x, F_y, F_z) and
// the readout in one kernel.  A workgroup walks a strip of kColStrip
// columns of kColX x kColY cells (all z) along y, one axis at a time.  For
// each column it transforms the z rows its particles' stencils touch,
// (kColX + 1) x (kColY + 1) of them, one wavefront per row, keeping the
// kColX + 1 rows it shares with the next column: the row's LDS slice serves
// as the transform's scratch and then holds the real row.  It then reads
// the force out at the column's particles.  The real force meshes are never
// written; a strip of 8 columns reads 15 + 7 x 10 spectrum rows for its 64
// owned ones (1.33 times each).  Particles come from the bins of a paint
// with kColX x kColY x tz tiles (z fastest, so a column's bins are one
// contiguous range).  The row transforms are the z c2r pass's arithmetic
// and the interpolation is k_tile_readout's, so forces are bitwise those of
// the unfused path.
//
// Software pipeline over the 3 x kColStrip (axis, column) units: each
// wavefront's first row of unit u + 1 is loaded into registers while unit
// u's particles are read out, and unit u's first record per thread is
// loaded before its rows are transformed, so neither load's latency stands
// alone between barriers.  Rows are dealt to the kColWaves wavefronts
// round-robin: a later column's 10 new rows take one each, the strip's
// first column (15 rows) gives three wavefronts a second row, loaded in
// place.  Measured at 512^3 particles / 1024^3 on the clustered run
// (readout ms per launch, 10-step average): 4 x 2 columns with 12
// wavefronts 6.2 (6.5 without the prefetch); 15 wavefronts 6.7-6.9, 10
// wavefronts 7.7; 2 x 4 columns 6.4-6.5; strips of 4 or 16 columns slower
// than 8.
constexpr int kColX = 4, kColY = 2;
constexpr int kColHY = kColY + 1;                 // y rows of a column: a ring of slices
constexpr int kColRows = (kColX + 1) * kColHY;    // row slices
constexpr int kColNew = (kColX + 1) * kColY;      // rows each later column of a strip adds
constexpr int kColStrip = 8;
constexpr int kColWaves = 12;
constexpr int kColUnits = 3 * kColStrip;
static_assert(kColWaves >= kColNew && 2 * kColWaves >= kColRows, "rows per wavefront");

__device__ __forceinline__ void column_row(const float2* s0, const float2* s1, const float2* s2,
                                           int u, int r, int x0, int Sy, int n, int cp,
                                           const float2*& src, int& slot) {
  const int a = u / kColStrip, c = u - a * kColStrip;
  int hx, hy;
  if (c == 0) {
    hx = r / kColHY;
    hy = r - hx * kColHY;
  } else {
    hx = r / kColY;
    hy = 1 + (r - hx * kColY);
  }
  int gx = x0 + hx, gy = (Sy * kColStrip + c) * kColY + hy;
  if (gx >= n) gx -= n;
  if (gy >= n) gy -= n;
  const float2* sa = s0;
  if (a == 1) sa = s1;
  if (a == 2) sa = s2;
  src = sa + ((int64_t)gx * n + gy) * cp;
  slot = hx * kColHY + (kColY * c + hy) % kColHY;
}

// The row's X[k] and X[M - k] for k = t + i T (X[M] for k = 0): one base per
// half and immediate offsets per element (the lane index is made opaque so
// that the compiler does not hoist 16 addresses out of the caller's loop).
template <int M, int T>
__device__ __forceinline__ void column_load_row(const float2* __restrict__ src, int t0,
                                                float2* pa, float2* pb) {
  int t = t0;
  asm volatile("" : "+v"(t));
  const float2* qa = src + t;
  const float2* qb = src + (M - t);
#pragma unroll
  for (int i = 0; i < M / T; ++i) {
    pa[i] = qa[i * T];
    pb[i] = qb[-i * T];
  }
}

// z c2r of one row from its loaded halves into LDS slice sp: the real row,
// x[2m] + i x[2m+1] = conj(z[m]), and x[n] = x[0] after it, so a stencil's
// z + 1 corner is always the next float.
template <int M, int T, class L>
__device__ __forceinline__ void column_transform_row(const float2* pa, const float2* pb,
                                                     float2* sp, int t,
                                                     const float2* stM, const float2* stN) {
  using namespace fftdev;
  float2 z[M / T];
#pragma unroll
  for (int i = 0; i < M / T; ++i) {
    const int k = t + i * T;
    z[i] = c2r_merge(pa[i], pb[i], stN[k], k == 0);
  }
  fft_regs<M, L, WaveSync>(z, sp, t, stM);
  WaveSync::sync();   // the row's lanes are done reading the last stage's input
#pragma unroll
  for (int i = 0; i < M / T; ++i) sp[t + i * T] = cconj(z[i]);
  if (t == 0) ((float*)sp)[2 * M] = z[0].x;
}

// Force component at one record from the column's real rows h (slices RS
// floats apart, ring of kColHY y rows per x row; the column's y0 row in
// ring slice ring0): k_tile_readout's interpolation, written to the
// particle's id-order slot.
template <int RS>
__device__ __forceinline__ void column_readout_one(const float* h, const float4 r4, float dx_inv,
                                                   int n, int x0, int y0, int ring0, float* fa) {
  static_assert(kColY + 1 == kColHY, "the stencil's y rows are two ring steps at most apart");
  const Cic1 cx = cic1(r4.x, dx_inv, n);
  const Cic1 cy = cic1(r4.y, dx_inv, n);
  const Cic1 cz = cic1(r4.z, dx_inv, n);
  const int lx = cx.i0 - x0;
  // ring slices of the stencil's y rows y0 + dy and y0 + dy + 1, dy < kColY
  int sa = ring0 + (cy.i0 - y0);
  if (sa >= kColHY) sa -= kColHY;
  int sb = sa + 1;
  if (sb >= kColHY) sb -= kColHY;
  const int q00 = (lx * kColHY + sa) * RS;
  const int q01 = (lx * kColHY + sb) * RS;
  const int q10 = q00 + kColHY * RS, q11 = q01 + kColHY * RS;
  const int z0 = cz.i0, z1 = cz.i0 + 1;
  // pm.cpp:203-218: f = 0; f += (ix0 terms); f += (ix1 terms)
  float v = 0.f;
  v += h[q00 + z0] * cx.w0 * cy.w0 * cz.w0 + h[q00 + z1] * cx.w0 * cy.w0 * cz.w1 +
       h[q01 + z0] * cx.w0 * cy.w1 * cz.w0 + h[q01 + z1] * cx.w0 * cy.w1 * cz.w1;
  v += h[q10 + z0] * cx.w1 * cy.w0 * cz.w0 + h[q10 + z1] * cx.w1 * cy.w0 * cz.w1 +
       h[q11 + z0] * cx.w1 * cy.w1 * cz.w0 + h[q11 + z1] * cx.w1 * cy.w1 * cz.w1;
  fa[__float_as_uint(r4.w)] = v;
}

// two workgroups per CU (LDS), 6 wavefronts per SIMD
template <int M>
__global__ void __launch_bounds__(kColWaves * 64) __attribute__((amdgpu_waves_per_eu(6)))
k_column_readout(const float2* __restrict__ s0, const float2* __restrict__ s1,
                 const float2* __restrict__ s2, const float4* __restrict__ rec,
                 const unsigned* __restrict__ offsets, TileGeom g,
                 const float2* __restrict__ twM, const float2* __restrict__ twN,
                 float* __restrict__ f, int64_t fstride) {
  using namespace fftdev;
  constexpr int T = Plan<M>::T;
  static_assert(T == 64, "one wavefront per z row");
  constexpr int I = M / T;         // elements per lane
  using L = ContigLayout;
  constexpr int LS = L::size(M);   // float2 per row slice: > M, a real row and one more float
  static_assert(LS > M, "row slice holds n + 1 floats");
  constexpr int RS = 2 * LS;       // floats per row slice
  __shared__ float2 rows[kColRows * LS];
  __shared__ float2 stM[M];
  __shared__ float2 stN[M];
  for (int i = threadIdx.x; i < M; i += blockDim.x) {
    stM[i] = twM[i];
    stN[i] = twN[i];
  }
  const int n = g.n;
  const int cp = g.ncz / 2;
  const float dx_inv = g.dx_inv;
  const unsigned nty = (unsigned)g.nty, ntz = (unsigned)g.ntz;
  // strips x fastest: the strips next to each other in x share their
  // boundary rows and run together on one XCD (its L2 serves the second read)
  const int nsx = n / kColX;
  const int b = xcd_remap(blockIdx.x, gridDim.x);
  const int Sy = b / nsx, Cx = b - Sy * nsx;
  const int x0 = Cx * kColX;
  // the wavefront index is uniform: its row bookkeeping stays scalar
  const int w0 = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), t0 = threadIdx.x & 63;
  const float* h = (const float*)rows;

  float2 pa[I], pb[I];
  const float2* src;
  int slot;
  column_row(s0, s1, s2, 0, w0, x0, Sy, n, cp, src, slot);
  column_load_row<M, T>(src, t0, pa, pb);
  __syncthreads();   // the twiddles are staged
#pragma unroll 1
  for (int u = 0; u < kColUnits; ++u) {
    const int a = u / kColStrip, c = u - a * kColStrip;
    const int Cy = Sy * kColStrip + c;
    const int y0 = Cy * kColY;
    const unsigned col = ((unsigned)Cx * nty + (unsigned)Cy) * ntz;
    const unsigned pbeg = offsets[col * 8u], pend = offsets[(col + ntz) * 8u];
    // this thread's first three records of the column, in flight during the
    // transform.  The loads are unconditional (record 0, which any record
    // scratch holds, past the column's end; not read out): a load behind a
    // branch is waited for where the branch ends, with every load before
    // it, the row prefetched for this unit's transform included.
    const unsigned j0 = pbeg + threadIdx.x, j1 = j0 + blockDim.x, j2 = j1 + blockDim.x;
    const float4 r0 = rec[j0 < pend ? j0 : 0u];
    const float4 r1 = rec[j1 < pend ? j1 : 0u];
    const float4 r2 = rec[j2 < pend ? j2 : 0u];
    const int nrows = c == 0 ? kColRows : kColNew;
    int r = w0;
    if (r < nrows) {   // the row loaded during the last unit
      // opaque copies of the lane and slice: without them the compiler
      // keeps every LDS address of the transform live across the loop
      int t = t0, sl = slot;
      asm volatile("" : "+v"(t), "+v"(sl));
      column_transform_row<M, T, L>(pa, pb, rows + sl * LS, t, stM, stN);
      r += kColWaves;
    }
    for (; r < nrows; r += kColWaves) {   // the first column's second rows
      column_row(s0, s1, s2, u, r, x0, Sy, n, cp, src, slot);
      column_load_row<M, T>(src, t0, pa, pb);
      int t = t0, sl = slot;
      asm volatile("" : "+v"(t), "+v"(sl));
      column_transform_row<M, T, L>(pa, pb, rows + sl * LS, t, stM, stN);
    }
    // the next unit's first row, in flight during this unit's readout
    if (u + 1 < kColUnits && w0 < ((u + 1) % kColStrip == 0 ? kColRows : kColNew)) {
      column_row(s0, s1, s2, u + 1, w0, x0, Sy, n, cp, src, slot);
      column_load_row<M, T>(src, t0, pa, pb);
    }
    __syncthreads();
    float* fa = f + a * fstride;
    const int ring0 = (kColY * c) % kColHY;
    if (j0 < pend) column_readout_one<RS>(h, r0, dx_inv, n, x0, y0, ring0, fa);
    if (j1 < pend) column_readout_one<RS>(h, r1, dx_inv, n, x0, y0, ring0, fa);
    if (j2 < pend) column_readout_one<RS>(h, r2, dx_inv, n, x0, y0, ring0, fa);
    for (unsigned j = j2 + blockDim.x; j < pend; j += blockDim.x)
      column_readout_one<RS>(h, rec[j], dx_inv, n, x0, y0, ring0, fa);
    __syncthreads();   // the rows are read before the next unit overwrites them
  }
}

// ---- adjoint kernels over the same bins ---------------------------------------------
// Max |v| of each of three value planes, as ordered float bits (|v| >= 0, so
// the unsigned order of the bits is the float order); amax cleared to 0.
// Grid-stride with a block reduction: one atomic per block and field.
__global__ void __launch_bounds__(kBlock)
k_absmax3(const float* __restrict__ v, int64_t stride, int64_t np, unsigned* __restrict__ amax) {
  __shared__ unsigned red[3][kBlock / 64];
  unsigned m[3] = {0u, 0u, 0u};
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < np;
       i += (int64_t)gridDim.x * blockDim.x) {
#pragma unroll
    for (int a = 0; a < 3; ++a) m[a] = max(m[a], __float_as_uint(fabsf(v[i + a * stride])));
  }
  const int wave = threadIdx.x >> 6;
#pragma unroll
  for (int a = 0; a < 3; ++a) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) m[a] = max(m[a], (unsigned)__shfl_xor((int)m[a], off));
    if (__lane_id() == 0) red[a][wave] = m[a];
  }
  __syncthreads();
  if (threadIdx.x < 3) {
    unsigned r = 0u;
#pragma unroll
    for (int w = 0; w < kBlock / 64; ++w) r = max(r, red[threadIdx.x][w]);
    if (r) atomicMax(amax + threadIdx.x, r);
  }
}

// Fixed-point scale of one value plane: a term w v (w <= 1) is bounded by
// amax < 2^e, so scale = 2^(30 - e) makes every term an int32 (one rounding,
// 2^(e - 31) at most) and the int64 cell sums exact.
__device__ __forceinline__ float value_scale(unsigned amax_bits) {
  const float amax = __uint_as_float(amax_bits);
  if (!(amax > 0.f)) return 1.f;
  if (!(amax < 3.0e38f)) return 0x1p-98f;   // inf / nan values: results are not finite anyway
  int e;
  frexpf(amax, &e);
  return ldexpf(1.f, 30 - e);
}

// Transpose of the three-field readout: out_a(c) = sum_p W(x_p - c) v_a(p),
// owner-computes per tile like k_tile_paint, with int64 fixed-point LDS sums
// (bitwise reproducible).  Every cell of the mesh is written.
__global__ void __launch_bounds__(1024)
k_tile_paint_values3(const float4* __restrict__ rec, const unsigned* __restrict__ offsets,
                     TileGeom g, const float* __restrict__ vals, int64_t vstride,
                     const unsigned* __restrict__ amax, float* __restrict__ o0,
                     float* __restrict__ o1, float* __restrict__ o2) {
  extern __shared__ unsigned char smem[];
  unsigned long long* acc = (unsigned long long*)smem;
  const int tcells = g.tx * g.ty * g.tz;
  lds_zero(smem, 3 * tcells * (int)sizeof(unsigned long long));

  // z-adjacent tiles (consecutive indices) on one XCD share halo lines in L2
  const TileCoord tc = tile_coord(g, xcd_remap(blockIdx.x, gridDim.x));
  const unsigned tile = (unsigned)tc.tile;
  offsets += (int64_t)tc.box * g.bins_box;
  const int64_t moff = (int64_t)tc.box * g.mstride;
  const int Tx = tc.Tx, Ty = tc.Ty, Tz = tc.Tz;
  const int x0 = Tx * g.tx, y0 = Ty * g.ty, z0 = Tz * g.tz;
  float sc[3];
#pragma unroll
  for (int a = 0; a < 3; ++a) sc[a] = value_scale(amax[a]);

  __shared__ unsigned rb[64], rpre[65];
  __shared__ int nranges;
  tile_source_ranges(offsets, g, Tx, Ty, Tz, rb, rpre, &nranges);
  const int nr = nranges;
  const unsigned total = rpre[nr];
  int r = 0;
  for (unsigned v = threadIdx.x; v < total; v += blockDim.x) {
    while (v >= rpre[r + 1]) ++r;
    const unsigned j = rb[r] + (v - rpre[r]);
    const float4 r4 = rec[j];
    const unsigned i = __float_as_uint(r4.w);
    float val[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) val[a] = vals[i + a * vstride];
    const Cic1 cx = cic1(r4.x, g.dx_inv, g.n);
    const Cic1 cy = cic1(r4.y, g.dx_inv, g.n);
    const Cic1 cz = cic1(r4.z, g.dx_inv, g.n);
    const int lx[2] = {cx.i0 - x0, cx.i1 - x0};
    const int ly[2] = {cy.i0 - y0, cy.i1 - y0};
    const int lz[2] = {cz.i0 - z0, cz.i1 - z0};
    const float wx[2] = {cx.w0, cx.w1};
    const float wy[2] = {cy.w0, cy.w1};
    const float wz[2] = {cz.w0, cz.w1};
#pragma unroll
    for (int ex = 0; ex < 2; ++ex) {
      if ((unsigned)lx[ex] >= (unsigned)g.tx) continue;
#pragma unroll
      for (int ey = 0; ey < 2; ++ey) {
        if ((unsigned)ly[ey] >= (unsigned)g.ty) continue;
#pragma unroll
        for (int ez = 0; ez < 2; ++ez) {
          if ((unsigned)lz[ez] >= (unsigned)g.tz) continue;
          const float w = wx[ex] * wy[ey] * wz[ez];
          const int li = (((lx[ex] << g.ly) + ly[ey]) << g.lz) + lz[ez];
#pragma unroll
          for (int a = 0; a < 3; ++a) {
            const int q = __float2int_rn(w * val[a] * sc[a]);
            atomicAdd(&acc[a * tcells + li], (unsigned long long)(long long)q);
          }
        }
      }
    }
  }
  __syncthreads();
  const int64_t t0 = moff + ((int64_t)x0 * g.n + y0) * g.ncz + z0;
  float* outs[3] = {o0 + t0, o1 + t0, o2 + t0};
  // sc is a power of two: multiplying by 1/sc is the exact division
  double isc[3];
#pragma unroll
  for (int a = 0; a < 3; ++a) isc[a] = 1.0 / (double)sc[a];
  for_tile_groups(g, [&](int c, int off, auto vc) {
    constexpr int V = decltype(vc)::value;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      float o[V];
#pragma unroll
      for (int k = 0; k < V; ++k) o[k] = (float)((double)(long long)acc[a * tcells + c + k] * isc[a]);
      store_group<V>(outs[a] + off, o);
    }
  });
}

// Gradient readout at the tile's own particles:
//   xbar_p += sum_c grad W(x_p - c) [fac rhobar(c) + sum_a fbar_a(p) F_a(c)]
// with rhobar (and, if FORCES, the three force meshes) staged over the tile
// and its +1 halo in LDS.  Same FP32 expression order as the gather kernel
// of adjoint_kernels.hip, so the two agree bitwise.
template <bool FORCES>
__global__ void __launch_bounds__(1024)
k_tile_grad_readout(const float4* __restrict__ rec, const unsigned* __restrict__ offsets,
                    TileGeom g, const float* __restrict__ rhobar, float fac,
                    const float* __restrict__ m0, const float* __restrict__ m1,
                    const float* __restrict__ m2, const float* __restrict__ fbar,
                    float* __restrict__ xbar, int64_t stride) {
  extern __shared__ float hs[];
  constexpr int NM = FORCES ? 4 : 1;
  const int hx = g.tx + 1, hy = g.ty + 1, hz = g.tz + 1;
  const int hcells = hx * hy * hz;
  // z-adjacent tiles (consecutive indices) on one XCD share halo lines in L2
  const TileCoord tc = tile_coord(g, xcd_remap(blockIdx.x, gridDim.x));
  const unsigned tile = (unsigned)tc.tile;
  offsets += (int64_t)tc.box * g.bins_box;
  const int64_t moff = (int64_t)tc.box * g.mstride;
  const int Tx = tc.Tx, Ty = tc.Ty, Tz = tc.Tz;
  const int x0 = Tx * g.tx, y0 = Ty * g.ty, z0 = Tz * g.tz;
  const float* ms[4] = {rhobar + moff, FORCES ? m0 + moff : nullptr, FORCES ? m1 + moff : nullptr,
                        FORCES ? m2 + moff : nullptr};
  stage_halo<NM>(hs, (int*)(hs + NM * hcells), ms, g, x0, y0, z0);
  const float dxi = g.dx_inv;
  const unsigned b = offsets[(unsigned)tile * 8u], e = offsets[(unsigned)tile * 8u + 8u];
  for (unsigned j = b + threadIdx.x; j < e; j += blockDim.x) {
    const float4 r4 = rec[j];
    const unsigned i = __float_as_uint(r4.w);
    const Cic1 cx = cic1(r4.x, dxi, g.n);
    const Cic1 cy = cic1(r4.y, dxi, g.n);
    const Cic1 cz = cic1(r4.z, dxi, g.n);
    float fb[3] = {0.f, 0.f, 0.f};
    if constexpr (FORCES) {
      fb[0] = fbar[i];
      fb[1] = fbar[i + stride];
      fb[2] = fbar[i + 2 * stride];
    }
    const int lx = cx.i0 - x0, ly = cy.i0 - y0, lz = cz.i0 - z0;
    float gx = 0.f, gy = 0.f, gz = 0.f;
#pragma unroll
    for (int ex = 0; ex < 2; ++ex)
#pragma unroll
      for (int ey = 0; ey < 2; ++ey)
#pragma unroll
        for (int ez = 0; ez < 2; ++ez) {
          const int c = ((lx + ex) * hy + (ly + ey)) * hz + (lz + ez);
          float S = fac * hs[c];
          if constexpr (FORCES)
            S += fb[0] * hs[hcells + c] + fb[1] * hs[2 * hcells + c] + fb[2] * hs[3 * hcells + c];
          const float wx = ex ? cx.w1 : cx.w0, wy = ey ? cy.w1 : cy.w0, wz = ez ? cz.w1 : cz.w0;
          const float sx = ex ? dxi : -dxi, sy = ey ? dxi : -dxi, sz = ez ? dxi : -dxi;
          gx += sx * wy * wz * S;
          gy += wx * sy * wz * S;
          gz += wx * wy * sz * S;
        }
    xbar[i] += gx;
    xbar[i + stride] += gy;
    xbar[i + 2 * stride] += gz;
  }
}

TileGeom make_geom(int n, float boxsize, const TileOpts& o, int64_t np = 0) {
  TileGeom g;
  g.n = n;
  g.ncz = grid_rpitch(n);
  // default 8 x 8 x 64 cells (measured fastest at 1024^3 for bin + scatter +
  // paint + readout): 256-byte rows along z, a 63 KB readout halo tile.
  // Meshes that 8 or 64 does not divide take the largest power of two that
  // divides n along each axis.
  auto pow2_divisor = [n](int cap) {
    int t = 1;
    while (t < cap && n % (2 * t) == 0) t *= 2;
    return t;
  };
  g.tx = o.tx ? o.tx : pow2_divisor(8);
  g.ty = o.ty ? o.ty : pow2_divisor(8);
  g.tz = o.tz ? o.tz : pow2_divisor(64);
  if ((g.tx & (g.tx - 1)) || (g.ty & (g.ty - 1)) || (g.tz & (g.tz - 1)) || g.tx < 1 ||
      g.ty < 1 || g.tz < 1)
    fail(FS2_ERR_VALUE, "paint tiles need power-of-two extents (n = %d, tile %dx%dx%d)", n, g.tx,
         g.ty, g.tz);
  if (n % g.tx || n % g.ty || n % g.tz)
    fail(FS2_ERR_VALUE, "tile %dx%dx%d does not divide the %d^3 mesh", g.tx, g.ty, g.tz, n);
  g.lx = __builtin_ctz((unsigned)g.tx);
  g.ly = __builtin_ctz((unsigned)g.ty);
  g.lz = __builtin_ctz((unsigned)g.tz);
  g.ntx = n / g.tx;
  g.nty = n / g.ty;
  g.ntz = n / g.tz;
  g.dx_inv = (float)n / boxsize;   // pm.cpp:77
  g.ntiles = g.ntx * g.nty * g.ntz;
  auto pow2 = [](int v) { return (v & (v - 1)) == 0; };
  if (pow2(g.ntx) && pow2(g.nty) && pow2(g.ntz)) {
    g.lntz = __builtin_ctz((unsigned)g.ntz);
    g.lnty = __builtin_ctz((unsigned)g.nty);
    g.lnt = __builtin_ctz((unsigned)g.ntiles);
    g.lg = std::min({kTileGroupLog, g.lnty, __builtin_ctz((unsigned)g.ntx)});
  } else {
    g.lntz = g.lnty = g.lnt = -1;
    g.lg = 0;
  }
  g.bins_box = (unsigned)g.ntiles * 8u;
  g.nbox = o.nbox < 1 ? 1 : o.nbox;
  if (np % g.nbox) fail(FS2_ERR_VALUE, "%lld particles do not split into %d boxes",
                        (long long)np, g.nbox);
  g.npb = (unsigned)(np / g.nbox);
  g.mstride = (int64_t)grid_floats(n);
  return g;
}

constexpr size_t kMaxLds = 160 * 1024;

void check_lds(size_t bytes, const char* what) {
  if (bytes > kMaxLds)
    fail(FS2_ERR_VALUE, "%s: tile needs %zu bytes of LDS (limit %zu)", what, bytes, kMaxLds);
}

void check_scratch(const TileGeom& g, int64_t np, const PaintScratch& s) {
  const int64_t nbins = (int64_t)g.nbox * g.bins_box;
  if ((size_t)nbins > s.nbins || (size_t)np > s.np)
    fail(FS2_ERR_RUNTIME, "paint scratch too small (%lld bins, %lld particles)",
         (long long)nbins, (long long)np);
}

}  // namespace

void tile_extents(int n, const TileOpts& o, int ext[3]) {
  const TileGeom g = make_geom(n, 1.0f, o);
  ext[0] = g.tx;
  ext[1] = g.ty;
  ext[2] = g.tz;
}

size_t paint_tiled_bins(int n, const TileOpts& o) {
  const TileGeom g = make_geom(n, 1.0f, o);
  return (size_t)g.nbox * g.bins_box;
}

namespace {
// K2 + K3 after the counts: the scan into offsets and cursors, the scatter.
void bin_scan_scatter(PosArgs p, const TileGeom& g, const PaintScratch& s) {
  const int64_t nbins = (int64_t)g.nbox * g.bins_box;
  const int nb = (int)((nbins + kScanChunk - 1) / kScanChunk);
  if (nb > kScanChunk) fail(FS2_ERR_VALUE, "too many paint bins (%lld)", (long long)nbins);
  // cursor holds the per-block partial scan until k_scan_add rewrites it
  launch("paint_scan", (double)nbins * 8.0, k_scan_blocks, dim3(nb), dim3(kScanThreads), 0,
         (const unsigned*)s.counts, nbins, s.cursor, s.block_sums);
  launch("paint_scan_sums", (double)nb * 8.0, k_scan_sums, dim3(1), dim3(kScanThreads), 0,
         s.block_sums, nb, s.offsets, nbins);
  launch("paint_scan_add", (double)nbins * 12.0, k_scan_add, dim3(blocks_for(nbins)),
         dim3(kBlock), 0, (const unsigned*)s.cursor, (const unsigned*)s.block_sums, nbins,
         s.offsets, s.cursor);
  if (p.np > 0) {
    launch("paint_scatter", (double)p.np * 28.0, k_scatter, dim3(bin_blocks(p.np)),
           dim3(kBinBlock), 0, (const float*)p.x, p.stride, p.np, g, s.cursor, s.rec);
  }
}

void clear_bins(const TileGeom& g, const PaintScratch& s) {
  const int64_t nbins = (int64_t)g.nbox * g.bins_box;
  launch("paint_clear_bins", (double)nbins * 4.0, k_fill_u32, dim3(blocks_for(nbins)),
         dim3(kBlock), 0, s.counts, nbins, 0u);
}
}  // namespace

void paint_tiled_bin(PosArgs p, int n, float boxsize, bool wrap, const PaintScratch& s,
                     const TileOpts& o) {
  const TileGeom g = make_geom(n, boxsize, o, p.np);
  check_scratch(g, p.np, s);
  clear_bins(g, s);
  if (p.np > 0) {
    launch("paint_bin", (double)p.np * 12.0, k_bin, dim3(bin_blocks(p.np)), dim3(kBinBlock), 0, p.x,
           p.stride, p.np, boxsize, wrap ? 1 : 0, g, s.counts);
  }
  bin_scan_scatter(p, g, s);
}

void paint_tiled_bin_kick_drift(PosArgs p, float* v, const float* f, const float* dx1,
                                const float* dx2, const KickDrift& kd, int n, float boxsize,
                                bool wrap, const PaintScratch& s, const TileOpts& o) {
  const TileGeom g = make_geom(n, boxsize, o, p.np);
  check_scratch(g, p.np, s);
  clear_bins(g, s);
  if (p.np > 0) {
    // the kick + drift pass's bytes (84 or 60 per particle); the binning
    // adds no traffic of its own
    launch("cola_kick_drift_bin", (double)p.np * (kd.kick ? 84.0 : 60.0), k_kick_drift_bin,
           dim3(bin_blocks(p.np)), dim3(kBinBlock), 0, p.x, v, f, dx1, dx2, p.stride, p.np, kd,
           boxsize, wrap ? 1 : 0, g, s.counts);
  }
  bin_scan_scatter(p, g, s);
}

namespace {
// fixed-point unit 2^-k with fac * 2^k <= 2^30, so every term fits an int32
struct FixedPoint {
  float fix;
  double fix_inv;
};
FixedPoint fixed_point(float fac) {
  int e = 0;
  std::frexp((double)std::max(fac, 1e-30f), &e);   // fac < 2^e
  const int k = std::min(60, 30 - e);
  return {(float)std::ldexp(1.0, k), std::ldexp(1.0, -k)};
}
}  // namespace

void paint_tiled_cells(int64_t np, float* grid, int n, float boxsize, float fac,
                       bool deterministic, const PaintScratch& s, const TileOpts& o) {
  const TileGeom g = make_geom(n, boxsize, o, np);
  check_scratch(g, np, s);
  const int64_t nblk = (int64_t)g.nbox * g.ntiles;
  const int tcells = g.tx * g.ty * g.tz;
  check_lds(tcells * (deterministic ? sizeof(unsigned long long) : sizeof(float)), "paint");
  const double bytes = (double)np * 12.0 + (double)n * n * n * 4.0 * g.nbox;
  const FixedPoint fp = fixed_point(fac);
  if (deterministic) {
    launch("paint_tile_det", bytes, k_tile_paint<unsigned long long>, dim3((unsigned)nblk),
           dim3(kTileBlock), tcells * sizeof(unsigned long long), (const float4*)s.rec,
           (const unsigned*)s.offsets, g, fac, fp.fix, fp.fix_inv, grid);
  } else {
    launch("paint_tile", bytes, k_tile_paint<float>, dim3((unsigned)nblk), dim3(kTileBlock),
           tcells * sizeof(float), (const float4*)s.rec, (const unsigned*)s.offsets,
           g, fac, fp.fix, fp.fix_inv, grid);
  }
}

void paint_tiled(PosArgs p, float* grid, int n, float boxsize, float fac, bool wrap,
                 bool deterministic, const PaintScratch& s, const TileOpts& o) {
  paint_tiled_bin(p, n, boxsize, wrap, s, o);
  paint_tiled_cells(p.np, grid, n, boxsize, fac, deterministic, s, o);
}

bool paint_r2c_supported(int n, const TileOpts& o) {
  return n == 1024 && o.tx * o.ty <= kR2cRowWaves && o.tz == n && o.nbox == 1;
}

void paint_tiled_r2c(int64_t np, float* grid, int n, float boxsize, float fac,
                     const PaintScratch& s, const TileOpts& o) {
  if (!paint_r2c_supported(n, o)) fail(FS2_ERR_VALUE, "fused paint + r2c: mesh %d", n);
  const TileGeom g = make_geom(n, boxsize, o, np);
  check_scratch(g, np, s);
  constexpr int M = 512;
  const int tcells = g.tx * g.ty * g.tz;
  check_lds(tcells * sizeof(unsigned long long) + 2 * M * sizeof(float2), "paint + r2c");
  // algorithmic bytes: positions read, the half spectrum written
  const double bytes = (double)np * 12.0 + (double)n * n * (M + 1) * 8.0;
  const FixedPoint fp = fixed_point(fac);
  if (g.ntiles % kR2cTiles)
    fail(FS2_ERR_VALUE, "fused paint + r2c: %d tiles are not a multiple of %d", g.ntiles,
         kR2cTiles);
  launch("paint_tile_r2c", bytes, k_tile_paint_r2c<M>, dim3((unsigned)(g.ntiles / kR2cTiles)),
         dim3(kTileR2cBlock),
         tcells * sizeof(unsigned long long), (const float4*)s.rec, (const unsigned*)s.offsets, g,
         fac, fp.fix, fp.fix_inv, fft_twiddle_table(M), fft_twiddle_table(2 * M), (float2*)grid);
}

void readout_tiled(const float* g0, const float* g1, const float* g2, float* f, int64_t fstride,
                   int64_t np, int n, float boxsize, const PaintScratch& s, const TileOpts& o) {
  const TileGeom g = make_geom(n, boxsize, o, np);
  check_scratch(g, np, s);
  const int64_t nblk = (int64_t)g.nbox * g.ntiles;
  const size_t lds = halo_lds(g, 3);
  check_lds(lds, "readout");
  const double bytes = (double)np * 24.0 + 3.0 * 4.0 * (double)n * n * n * g.nbox;
  launch("readout_tile", bytes, k_tile_readout, dim3((unsigned)nblk), dim3(kReadoutBlock), lds,
         (const float4*)s.rec, (const unsigned*)s.offsets, g, g0, g1, g2, f, fstride);
}

bool column_readout_supported(int n) { return n == 1024; }

TileOpts column_tile_opts(int n) {
  TileOpts o;
  if (column_readout_supported(n)) {   // full z rows: the paint can fuse the z r2c too
    o.tx = kColX;
    o.ty = kColY;
    o.tz = n;
  }
  return o;
}

void readout_columns(const float* s0, const float* s1, const float* s2, float* f,
                     int64_t fstride, int64_t np, int n, float boxsize, const PaintScratch& s,
                     const TileOpts& o) {
  if (!column_readout_supported(n)) fail(FS2_ERR_VALUE, "column readout: mesh %d", n);
  const TileGeom g = make_geom(n, boxsize, o, np);
  check_scratch(g, np, s);
  if (g.tx != kColX || g.ty != kColY || g.nbox != 1 || (n / kColY) % kColStrip)
    fail(FS2_ERR_RUNTIME, "column readout needs the bins of %dx%dx* paint tiles (have %dx%dx%d)",
         kColX, kColY, g.tx, g.ty, g.tz);
  constexpr int M = 512;
  const unsigned nstrip = (unsigned)(n / kColX) * (unsigned)(n / kColY / kColStrip);
  // algorithmic bytes: the three spectra read once, positions and forces
  const double bytes = 3.0 * (double)n * n * (M + 1) * 8.0 + (double)np * 24.0;
  launch("readout_columns", bytes, k_column_readout<M>, dim3(nstrip), dim3(kColWaves * 64), 0,
         (const float2*)s0, (const float2*)s1, (const float2*)s2, (const float4*)s.rec,
         (const unsigned*)s.offsets, g, fft_twiddle_table(M), fft_twiddle_table(2 * M), f,
         fstride);
}

void paint_values3_tiled(const float* vals, int64_t vstride, int64_t np, float* const out[3],
                         int n, float boxsize, const PaintScratch& s, const TileOpts& o) {
  const TileGeom g = make_geom(n, boxsize, o, np);
  check_scratch(g, np, s);
  const int64_t nblk = (int64_t)g.nbox * g.ntiles;
  const int tcells = g.tx * g.ty * g.tz;
  check_lds(3 * tcells * sizeof(unsigned long long), "adjoint paint");
  launch("adj_absmax_clear", 12.0, k_fill_u32, dim3(1), dim3(kBlock), 0, s.amax, (int64_t)3, 0u);
  if (np > 0)
    launch("adj_absmax3", (double)np * 12.0, k_absmax3,
           dim3(std::min<unsigned>(blocks_for(np), 1024u)), dim3(kBlock), 0, vals, vstride, np,
           s.amax);
  launch("adj_paint_tile3", (double)np * 28.0 + 12.0 * (double)n * n * n * g.nbox,
         k_tile_paint_values3, dim3((unsigned)nblk), dim3(1024),
         3 * tcells * sizeof(unsigned long long), (const float4*)s.rec,
         (const unsigned*)s.offsets, g, vals, vstride, (const unsigned*)s.amax, out[0],
         out[1], out[2]);
}

void grad_readout_tiled(const float* rhobar, float fac, const float* const F[3],
                        const float* fbar, float* xbar, int64_t stride, int64_t np, int n,
                        float boxsize, const PaintScratch& s, const TileOpts& o) {
  const TileGeom g = make_geom(n, boxsize, o, np);
  check_scratch(g, np, s);
  const int64_t nblk = (int64_t)g.nbox * g.ntiles;
  check_lds(halo_lds(g, F ? 4 : 1), "gradient readout");
  if (F) {
    launch("adj_grad_tile4", (double)np * 40.0 + 16.0 * (double)n * n * n * g.nbox,
           k_tile_grad_readout<true>, dim3((unsigned)nblk), dim3(1024), halo_lds(g, 4),
           (const float4*)s.rec, (const unsigned*)s.offsets, g, rhobar, fac, F[0], F[1], F[2],
           fbar, xbar, stride);
  } else {
    launch("adj_grad_tile1", (double)np * 28.0 + 4.0 * (double)n * n * n * g.nbox,
           k_tile_grad_readout<false>, dim3((unsigned)nblk), dim3(1024), halo_lds(g, 1),
           (const float4*)s.rec, (const unsigned*)s.offsets, g, rhobar, fac, nullptr, nullptr,
           nullptr, nullptr, xbar, stride);
  }
}

}  // namespace fs2
