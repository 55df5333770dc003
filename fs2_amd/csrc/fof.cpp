// Friends-of-friends halo finder on the host (py/fof.cpp:45-161 with the
// kd-tree of py/kdtree.cpp:70-215).  The positions arrive from the device
// through a pinned buffer (capi.cpp: fs2_fof_find_groups); this file only
// sees plain host arrays.
//
// Observable results kept from the reference:
//  - the particle order.  The reference builds its tree by recursive median
//    splits (std::nth_element) of the particle array itself, so FoF leaves
//    the particles permuted.  The same splits run here on (x, index)
//    records and the permutation is returned (new -> old).  A node of
//    n <= quota particles is a leaf; the cut axis is the longest edge of a
//    cuboid that starts as boxsize3 (or the bounding box) and is halved
//    along the cut at every split.  The reference halves one shared array
//    in depth-first order, so a right subtree sees the halvings of the left
//    one (and after a few hundred splits the cuboid is all zeros and every
//    cut is along y); that sequence depends only on the node sizes, so it
//    is replayed first (plan) and the subtrees are then split in parallel.
//  - the tree copy: each node stores the [left, right] extent of its
//    particles along its parent's cut axis (the root: along x).
//  - the groups: i and j are friends when the FP32 periodic distance
//    (dx wrapped once into [-L/2, L/2), d2 = dx^2 + dy^2 + dz^2) is below
//    ll^2, and grp[i] is the lowest index (in the new order) of i's
//    connected set of friends, which is what the reference's breadth-first
//    growth from the lowest unassigned index yields.
// The friends are found here on a periodic cell grid (cells at least ll
// wide) with a concurrent union-find, not by the reference's tree walk:
// the walk prunes along one axis per node, which the all-y cuts above make
// quadratic-like at large n, and its FP32 pruning test can skip a pair
// whose distance is within an ulp of ll.  The grid examines every pair
// that can be friends, so the groups are exactly those of the FP32 pair
// criterion.
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <exception>
#include <system_error>
#include <thread>
#include <vector>

#include "fs2_internal.h"

#pragma clang fp contract(off)

namespace fs2 {
namespace {

struct Rec {
  float x[3];
  int32_t idx;   // index before the tree build
};

int host_threads() {
  int n = (int)std::thread::hardware_concurrency();
  if (const char* s = std::getenv("OMP_NUM_THREADS")) {
    const int e = std::atoi(s);
    if (e > 0) n = e;
  }
  return std::max(1, std::min(n, 64));
}

// Runs f(t) for t < nt on nt threads (the caller's thread takes t = 0, and
// any t whose thread cannot be started).  Every thread is joined before an
// exception of any f(t) is rethrown on the caller's thread.
template <typename F>
void parallel(int nt, F f) {
  std::vector<std::exception_ptr> err((size_t)std::max(nt, 1));
  auto run = [&](int t) {
    try {
      f(t);
    } catch (...) {
      err[(size_t)t] = std::current_exception();
    }
  };
  std::vector<std::thread> th;
  th.reserve((size_t)std::max(nt - 1, 0));
  for (int t = 1; t < nt; ++t) {
    try {
      th.emplace_back(run, t);
    } catch (const std::system_error&) {
      run(t);
    }
  }
  run(0);
  for (auto& x : th) x.join();
  for (auto& e : err)
    if (e) std::rethrow_exception(e);
}

struct Builder {
  std::vector<Rec>& v;
  std::vector<KdNode>& tree;
  int quota;
  std::vector<int8_t> cut;   // cut axis of every internal node

  static int cut_direction(const float b[3]) {   // longest edge, ties as kdtree.cpp:49-59
    int k = 1;
    if (b[2] > b[1]) k = 2;
    if (b[0] > b[k]) k = 0;
    return k;
  }

  // The reference's depth-first sequence of halvings of one cuboid.
  void plan(size_t inode, int32_t ib, int32_t ie, float box[3]) {
    if (ie - ib <= quota) return;
    const int k = cut_direction(box);
    box[k] /= 2;
    cut[inode] = (int8_t)k;
    const int32_t mid = ib + (ie - ib) / 2;
    plan(2 * inode + 1, ib, mid, box);
    plan(2 * inode + 2, mid, ie, box);
  }

  void bounding_box(int32_t ib, int32_t ie, float lo[3], float hi[3]) const {
    for (int a = 0; a < 3; ++a) lo[a] = hi[a] = v[ib].x[a];
    for (int32_t i = ib + 1; i < ie; ++i)
      for (int a = 0; a < 3; ++a) {
        lo[a] = std::min(lo[a], v[i].x[a]);
        hi[a] = std::max(hi[a], v[i].x[a]);
      }
  }

  // Splits node inode over [ib, ie) and returns its bounding box in lo/hi;
  // the two subtrees of the top `spawn` levels run on separate threads.
  void build(size_t inode, int32_t ib, int32_t ie, float lo[3], float hi[3], int spawn) {
    KdNode& t = tree[inode];
    t.ibegin = ib;
    t.iend = ie;
    if (ie - ib <= quota) {
      bounding_box(ib, ie, lo, hi);
      t.left = lo[t.k];
      t.right = hi[t.k];
      return;
    }
    const int k = cut[inode];
    const int32_t mid = ib + (ie - ib) / 2;
    std::nth_element(v.begin() + ib, v.begin() + mid, v.begin() + ie,
                     [k](const Rec& a, const Rec& b) { return a.x[k] < b.x[k]; });
    const size_t il = 2 * inode + 1, ir = 2 * inode + 2;
    tree[il].k = k;
    tree[ir].k = k;
    float lo1[3], hi1[3];
    std::thread th;
    if (spawn > 0) {
      try {
        th = std::thread([&] { build(il, ib, mid, lo, hi, spawn - 1); });
      } catch (const std::system_error&) {   // no thread: build both here
      }
    }
    if (th.joinable()) {
      build(ir, mid, ie, lo1, hi1, spawn - 1);   // build never throws
      th.join();
    } else {
      build(il, ib, mid, lo, hi, 0);
      build(ir, mid, ie, lo1, hi1, 0);
    }
    tree[il].left = lo[k];
    tree[il].right = hi[k];
    tree[ir].left = lo1[k];
    tree[ir].right = hi1[k];
    for (int a = 0; a < 3; ++a) {
      lo[a] = std::min(lo[a], lo1[a]);
      hi[a] = std::max(hi[a], hi1[a]);
    }
  }
};

// Concurrent union-find whose roots are always the lowest index of their
// set: a root is only ever attached below a smaller root.
struct MinUnionFind {
  std::vector<std::atomic<int32_t>> parent;
  explicit MinUnionFind(int64_t n) : parent((size_t)n) {
    for (int64_t i = 0; i < n; ++i) parent[i].store((int32_t)i, std::memory_order_relaxed);
  }
  int32_t find(int32_t x) {
    for (;;) {
      int32_t p = parent[x].load(std::memory_order_relaxed);
      if (p == x) return x;
      const int32_t g = parent[p].load(std::memory_order_relaxed);
      if (g != p) parent[x].compare_exchange_weak(p, g, std::memory_order_relaxed);   // halving
      x = g;
    }
  }
  void unite(int32_t a, int32_t b) {
    for (;;) {
      a = find(a);
      b = find(b);
      if (a == b) return;
      if (a > b) std::swap(a, b);
      int32_t expect = b;
      if (parent[b].compare_exchange_strong(expect, a, std::memory_order_relaxed)) return;
    }
  }
};

struct PairTest {
  float L, half, ll2;
  float periodic(float d) const {   // fof.cpp:21-27
    d = d < -half ? d + L : d;
    d = d >= half ? d - L : d;
    return d;
  }
  bool friends(const float* a, const float* b) const {   // fof.cpp:29-37
    const float dx = periodic(a[0] - b[0]);
    const float dy = periodic(a[1] - b[1]);
    const float dz = periodic(a[2] - b[2]);
    return dx * dx + dy * dy + dz * dz < ll2;
  }
};

// Friends of every particle of v on a periodic grid of nc^3 cells of width
// L / nc >= ll (a friend is at most one cell away along each axis, also for
// positions outside [0, L): the one-wrap distance is small only if the
// wrapped positions are close).  Half-shell pairs, one task per x column
// of cells.
void link_on_grid(const std::vector<Rec>& v, double L, float ll, const PairTest& pt,
                  MinUnionFind& uf, int nt) {
  const int64_t n = (int64_t)v.size();
  if (!(pt.ll2 > 0.f)) return;   // d2 < ll^2 <= 0 never holds
  int nc = 1;
  {
    const double cells = std::floor(L / (double)ll);
    nc = (int)std::max(1.0, std::min({cells, std::cbrt((double)n), 1024.0}));
  }
  if (nc < 3) nc = 1;   // every pair in one cell
  const int64_t ncell = (int64_t)nc * nc * nc;
  auto cell_of = [&](const float* x) {
    int64_t c = 0;
    for (int a = 0; a < 3; ++a) {
      double u = std::fmod((double)x[a], L);
      if (u < 0) u += L;
      int ci = (int)(u / L * nc);
      ci = std::min(std::max(ci, 0), nc - 1);
      c = c * nc + ci;
    }
    return c;
  };
  std::vector<int64_t> cell((size_t)n);
  std::vector<int64_t> start((size_t)ncell + 1, 0);
  for (int64_t i = 0; i < n; ++i) {
    cell[i] = cell_of(v[i].x);
    ++start[cell[i] + 1];
  }
  for (int64_t c = 0; c < ncell; ++c) start[c + 1] += start[c];
  std::vector<int32_t> order((size_t)n);
  {
    std::vector<int64_t> cur(start.begin(), start.end() - 1);
    for (int64_t i = 0; i < n; ++i) order[cur[cell[i]]++] = (int32_t)i;
  }
  std::vector<int64_t>().swap(cell);

  // the 13 neighbours after this cell in (dx, dy, dz) order
  int off[13][3];
  int no = 0;
  for (int dx = -1; dx <= 1; ++dx)
    for (int dy = -1; dy <= 1; ++dy)
      for (int dz = -1; dz <= 1; ++dz)
        if (dx > 0 || (dx == 0 && (dy > 0 || (dy == 0 && dz > 0)))) {
          off[no][0] = dx;
          off[no][1] = dy;
          off[no][2] = dz;
          ++no;
        }
  std::atomic<int> next{0};
  parallel(nt, [&](int) {
    for (int cx; (cx = next.fetch_add(1)) < nc;) {
      for (int cy = 0; cy < nc; ++cy)
        for (int cz = 0; cz < nc; ++cz) {
          const int64_t c = ((int64_t)cx * nc + cy) * nc + cz;
          const int64_t b0 = start[c], e0 = start[c + 1];
          for (int64_t s = b0; s < e0; ++s) {
            const int32_t i = order[s];
            for (int64_t t = s + 1; t < e0; ++t)
              if (pt.friends(v[i].x, v[order[t]].x)) uf.unite(i, order[t]);
          }
          if (nc == 1 || b0 == e0) continue;
          for (int o = 0; o < 13; ++o) {
            const int nx = (cx + off[o][0] + nc) % nc;
            const int ny = (cy + off[o][1] + nc) % nc;
            const int nz = (cz + off[o][2] + nc) % nc;
            const int64_t d = ((int64_t)nx * nc + ny) * nc + nz;
            const int64_t b1 = start[d], e1 = start[d + 1];
            for (int64_t s = b0; s < e0; ++s) {
              const int32_t i = order[s];
              for (int64_t t = b1; t < e1; ++t)
                if (pt.friends(v[i].x, v[order[t]].x)) uf.unite(i, order[t]);
            }
          }
        }
    }
  });
}

}  // namespace

void fof_groups(const float* const xyz[3], int64_t n, int64_t np_alloc, double boxsize,
                double linking_length, const double* boxsize3, int quota, FofResult& out) {
  if (quota < 1) fail(FS2_ERR_VALUE, "kd-tree quota must be positive (got %d)", quota);
  if (n > INT32_MAX || np_alloc > INT32_MAX)
    fail(FS2_ERR_VALUE, "FoF indexes particles with int32 (n = %lld)", (long long)n);
  // tree size from the allocation, kdtree.cpp:131-140
  int64_t nleaf = 1;
  int height = 0;
  while ((int64_t)quota * nleaf < std::max<int64_t>(np_alloc, n)) {
    nleaf <<= 1;
    ++height;
  }
  out.height = height;
  out.quota = quota;
  out.tree.assign((size_t)(2 * nleaf - 1), KdNode{0, 0.f, 0.f, 0, 0});
  out.grp.resize((size_t)n);
  out.perm.resize((size_t)n);
  if (n == 0) return;
  const int nt = host_threads();

  std::vector<Rec> v((size_t)n);
  for (int64_t i = 0; i < n; ++i) {
    for (int a = 0; a < 3; ++a) v[i].x[a] = xyz[a][i];
    v[i].idx = (int32_t)i;
  }
  float lo[3], hi[3], box[3];
  Builder b{v, out.tree, quota, std::vector<int8_t>(out.tree.size(), 0)};
  if (boxsize3) {
    for (int a = 0; a < 3; ++a) box[a] = (float)boxsize3[a];
  } else {
    b.bounding_box(0, (int32_t)n, lo, hi);
    for (int a = 0; a < 3; ++a) box[a] = hi[a] - lo[a];
  }
  b.plan(0, 0, (int32_t)n, box);
  int spawn = 0;
  while ((1 << spawn) < nt && spawn < 6) ++spawn;
  b.build(0, 0, (int32_t)n, lo, hi, spawn);
  out.tree[0].k = 0;
  out.tree[0].left = lo[0];
  out.tree[0].right = hi[0];
  for (int64_t i = 0; i < n; ++i) out.perm[i] = v[i].idx;

  const float ll = (float)linking_length;
  const PairTest pt{(float)boxsize, (float)(boxsize / 2), ll * ll};
  MinUnionFind uf(n);
  link_on_grid(v, boxsize, ll, pt, uf, nt);
  parallel(nt, [&](int t) {
    for (int64_t i = t; i < n; i += nt) out.grp[i] = uf.find((int32_t)i);
  });
}

void fof_counts(const std::vector<int32_t>& grp, std::vector<int32_t>& nfof) {
  nfof.assign(grp.size(), 0);   // fof.cpp:101-113
  for (int32_t g : grp) nfof[g]++;
}

}  // namespace fs2
