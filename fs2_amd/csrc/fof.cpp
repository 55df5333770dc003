// Friends-of-friends halo finder over a balanced kd-tree, on the host
// (py/fof.cpp:45-161, py/kdtree.cpp:70-215).  The positions arrive from the
// device through a pinned buffer (capi.cpp: fs2_fof_find_groups); this file
// only sees plain host arrays.
//
// Semantics kept from the reference, because they are observable:
//  - the tree is built by recursive median splits (std::nth_element) of the
//    particle array itself, so FoF leaves the particles permuted; the
//    permutation is returned (new -> old) and applied to the device planes;
//  - a node of n <= quota particles is a leaf; the cut direction is the
//    longest edge of a cuboid that starts as boxsize3 (or the bounding box)
//    and is halved along the cut at every level; each node stores the
//    [left, right] extent of its particles along its parent's cut axis
//    (the root: along x);
//  - groups are grown breadth-first from the lowest unassigned index, so
//    grp[i] is the lowest (permuted) index of i's group;
//  - distances are periodic and computed in FP32 like the reference:
//    dx wrapped once into [-L/2, L/2), d2 = dx^2 + dy^2 + dz^2 < ll^2, with
//    the same FP32 node pruning test.
#include <algorithm>
#include <cstdint>
#include <queue>
#include <vector>

#include "fs2_internal.h"

#pragma clang fp contract(off)

namespace fs2 {
namespace {

struct Rec {
  float x[3];
  int32_t idx;   // index before the tree build
};

struct Builder {
  std::vector<Rec>& v;
  std::vector<KdNode>& tree;
  int quota;

  static int cut_direction(const float b[3]) {   // longest edge, ties as kdtree.cpp:49-59
    int k = 1;
    if (b[2] > b[1]) k = 2;
    if (b[0] > b[k]) k = 0;
    return k;
  }

  void bounding_box(int32_t ib, int32_t ie, float lo[3], float hi[3]) const {
    for (int a = 0; a < 3; ++a) lo[a] = hi[a] = v[ib].x[a];
    for (int32_t i = ib + 1; i < ie; ++i)
      for (int a = 0; a < 3; ++a) {
        lo[a] = std::min(lo[a], v[i].x[a]);
        hi[a] = std::max(hi[a], v[i].x[a]);
      }
  }

  // Builds node inode over [ib, ie) and returns its bounding box in lo/hi.
  void build(size_t inode, int32_t ib, int32_t ie, float lo[3], float hi[3], float box[3]) {
    KdNode& t = tree[inode];
    t.ibegin = ib;
    t.iend = ie;
    if (ie - ib <= quota) {
      bounding_box(ib, ie, lo, hi);
      t.left = lo[t.k];
      t.right = hi[t.k];
      return;
    }
    const int k = cut_direction(box);
    box[k] /= 2;
    const int32_t mid = ib + (ie - ib) / 2;
    std::nth_element(v.begin() + ib, v.begin() + mid, v.begin() + ie,
                     [k](const Rec& a, const Rec& b) { return a.x[k] < b.x[k]; });
    const size_t il = 2 * inode + 1, ir = 2 * inode + 2;
    tree[il].k = k;
    build(il, ib, mid, lo, hi, box);
    tree[il].left = lo[k];
    tree[il].right = hi[k];
    float lo1[3], hi1[3];
    tree[ir].k = k;
    build(ir, mid, ie, lo1, hi1, box);
    tree[ir].left = lo1[k];
    tree[ir].right = hi1[k];
    for (int a = 0; a < 3; ++a) {
      lo[a] = std::min(lo[a], lo1[a]);
      hi[a] = std::max(hi[a], hi1[a]);
    }
  }
};

struct Linker {
  const std::vector<Rec>& v;
  const std::vector<KdNode>& tree;
  std::vector<int32_t>& grp;
  std::queue<int32_t>& q;
  int quota;
  float L, half, ll, ll2;

  float periodic(float d) const {
    d = d < -half ? d + L : d;
    d = d >= half ? d - L : d;
    return d;
  }
  float dist2(const float* a, const float* b) const {
    const float dx = periodic(a[0] - b[0]);
    const float dy = periodic(a[1] - b[1]);
    const float dz = periodic(a[2] - b[2]);
    return dx * dx + dy * dy + dz * dz;
  }

  // Queues the unlinked friends of particle i found under node inode.
  void link(size_t inode, int32_t i) {
    const KdNode& t = tree[inode];
    const float x = v[i].x[t.k];
    if ((x < t.left - ll && x + L > t.right + ll) || (x > t.right + ll && x - L < t.left - ll))
      return;   // node out of reach along its axis, also across the boundary
    if (t.iend - t.ibegin <= quota) {
      for (int32_t j = t.ibegin; j < t.iend; ++j) {
        if (grp[j] != grp[i] && dist2(v[i].x, v[j].x) < ll2) {
          q.push(j);
          grp[j] = grp[i];
        }
      }
      return;
    }
    link(2 * inode + 1, i);
    link(2 * inode + 2, i);
  }
};

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

  std::vector<Rec> v((size_t)n);
  for (int64_t i = 0; i < n; ++i) {
    for (int a = 0; a < 3; ++a) v[i].x[a] = xyz[a][i];
    v[i].idx = (int32_t)i;
  }
  float lo[3], hi[3], box[3];
  Builder b{v, out.tree, quota};
  if (boxsize3) {
    for (int a = 0; a < 3; ++a) box[a] = (float)boxsize3[a];
  } else {
    b.bounding_box(0, (int32_t)n, lo, hi);
    for (int a = 0; a < 3; ++a) box[a] = hi[a] - lo[a];
  }
  b.build(0, 0, (int32_t)n, lo, hi, box);
  out.tree[0].k = 0;
  out.tree[0].left = lo[0];
  out.tree[0].right = hi[0];

  for (int64_t i = 0; i < n; ++i) {
    out.perm[i] = v[i].idx;
    out.grp[i] = (int32_t)i;
  }
  std::queue<int32_t> q;
  const float L = (float)boxsize;
  const float ll = (float)linking_length;
  Linker lk{v, out.tree, out.grp, q, quota, L, (float)(boxsize / 2), ll, ll * ll};
  for (int32_t i = 0; i < (int32_t)n; ++i) {
    if (out.grp[i] != i) continue;   // already in an earlier group
    q.push(i);
    while (!q.empty()) {
      lk.link(0, q.front());
      q.pop();
    }
  }
}

void fof_counts(const std::vector<int32_t>& grp, std::vector<int32_t>& nfof) {
  nfof.assign(grp.size(), 0);   // fof.cpp:101-113
  for (int32_t g : grp) nfof[g]++;
}

}  // namespace fs2
