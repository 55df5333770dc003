"""Friends-of-friends oracle (test infrastructure only).

Restates py/fof.cpp:45-161 and the layout of py/kdtree.cpp:94-190 for the
checks of fs2_fof_find_groups:

- ``groups``: brute force over all pairs (FP32 periodic distances as
  fof.cpp:24-39: dx wrapped once into [-L/2, L/2), d2 < ll^2), union-find,
  every particle labelled with the lowest index of its group.  The
  reference's breadth-first growth from the lowest unassigned index gives
  exactly these labels when the friend relation is symmetric, which the
  FP32 distance is (x - y = -(y - x) exactly).
- ``tree_layout``: node ranges, cut axes and the member set of every node.
  With distinct coordinates nth_element's partition fixes the set of each
  child (the lower half by the cut coordinate), so the sets, ranges, axes
  and [left, right] extents are implementation independent; only the order
  inside a leaf is not.  The cut cuboid is one array halved in depth-first
  order, as kdtree.cpp:96-105 does (the right subtree sees the halvings of
  the left one).

Known answers pinned in tests/test_oracle_fof.py: the line cases of the
reference's test/test_fof.py:40-74 and closed-form small configurations.
"""
import numpy as np


def periodic_d2(a, b, L):
    """FP32 squared periodic distance between rows of a and b (broadcast)."""
    L = np.float32(L)
    half = np.float32(L / np.float32(2))
    d = (a.astype(np.float32)[:, None, :] - b.astype(np.float32)[None, :, :]).astype(np.float32)
    d = np.where(d < -half, d + L, d).astype(np.float32)
    d = np.where(d >= half, d - L, d).astype(np.float32)
    return (d[..., 0] * d[..., 0] + d[..., 1] * d[..., 1] + d[..., 2] * d[..., 2]).astype(np.float32)


def groups(x, L, ll, chunk=1024):
    """grp[i] = lowest index of i's group (x: (n, 3) float32, in the order
    the finder sees them)."""
    x = np.asarray(x, dtype=np.float32)
    n = len(x)
    ll2 = np.float32(ll) * np.float32(ll)
    parent = np.arange(n)

    def find(i):
        while parent[i] != i:
            parent[i] = parent[parent[i]]
            i = parent[i]
        return i

    for s in range(0, n, chunk):
        d2 = periodic_d2(x[s:s + chunk], x, L)
        ii, jj = np.nonzero(d2 < ll2)
        for i, j in zip(ii + s, jj):
            if i < j:
                ri, rj = find(i), find(j)
                if ri != rj:
                    parent[max(ri, rj)] = min(ri, rj)
    roots = np.array([find(i) for i in range(n)], dtype=np.int64)
    # root = lowest index: union always keeps the smaller root
    return roots.astype(np.int32)


def nfof(grp):
    out = np.zeros(len(grp), dtype=np.int32)
    np.add.at(out, grp, 1)
    return out


def tree_height(np_alloc, quota):
    """kdtree.cpp:131-137: smallest h with quota 2^h >= np_alloc."""
    nleaf, h = 1, 0
    while quota * nleaf < np_alloc:
        nleaf <<= 1
        h += 1
    return h


def tree_layout(x, quota, boxsize3=None):
    """Nodes of the balanced kd-tree over x (distinct coordinates):
    {inode: (k, left, right, ibegin, iend, member_set)} for the built nodes.
    x is indexed by the particle order before the build."""
    x = np.asarray(x, dtype=np.float32)
    n = len(x)
    if boxsize3 is None:
        box = (x.max(axis=0) - x.min(axis=0)).astype(np.float32)
    else:
        box = np.asarray(boxsize3, dtype=np.float32).copy()
    nodes = {}

    def cut(b):
        k = 1
        if b[2] > b[1]:
            k = 2
        if b[0] > b[k]:
            k = 0
        return k

    def build(inode, ib, ie, members, k_parent):
        members = np.asarray(members)
        if ie - ib <= quota:
            nodes[inode] = [k_parent, ib, ie, members]
            return
        k = cut(box)
        box[k] = np.float32(box[k] / np.float32(2))
        mid = ib + (ie - ib) // 2
        order = members[np.argsort(x[members, k], kind="stable")]
        nodes[inode] = [k_parent, ib, ie, members]
        build(2 * inode + 1, ib, mid, order[:mid - ib], k)
        build(2 * inode + 2, mid, ie, order[mid - ib:], k)

    build(0, 0, n, np.arange(n), 0)
    out = {}
    for i, (k, ib, ie, m) in nodes.items():
        out[i] = (k, float(x[m, k].min()) if len(m) else 0.0,
                  float(x[m, k].max()) if len(m) else 0.0, ib, ie, frozenset(m.tolist()))
    return out
