"""The FoF oracle against known answers: the line cases of the reference's
test/test_fof.py:40-74, periodic linking across the boundary, a literal
restatement of the breadth-first growth of py/fof.cpp:78-96 without a tree,
and a hand-derived kd-tree layout (py/kdtree.cpp:94-125)."""
from collections import deque

import numpy as np
import pytest

from oracle import fof as ofof


def _line(x0, dx, n):
    return np.array([[x0[a] + i * dx[a] for a in range(3)] for i in range(n)], dtype=np.float32)


# test/test_fof.py:11-16: nc = 4, boxsize = 1, dx = boxsize / nc, ll = 0.2 dx
NC, BOX = 4, 1.0
DX = BOX / NC
LL = 0.2 * BOX / NC


def test_line_sparse_is_n_groups():
    x = _line([0.5 * DX] * 3, [1.1 * LL, 0, 0], 10)
    grp = ofof.groups(x, BOX, LL)
    assert np.array_equal(grp, np.arange(10))
    assert np.count_nonzero(ofof.nfof(grp)) == 10


def test_line_dense_is_one_group():
    x = _line([0.5 * DX] * 3, [0.99 * LL, 0, 0], 10)
    grp = ofof.groups(x, BOX, LL)
    assert np.all(grp == 0)
    n = ofof.nfof(grp)
    assert n[0] == 10 and np.count_nonzero(n) == 1


def test_links_across_the_periodic_boundary():
    x = np.array([[0.01, 0.5, 0.5], [0.3, 0.3, 0.3], [BOX - 0.02, 0.5, 0.5]], dtype=np.float32)
    assert list(ofof.groups(x, BOX, LL)) == [0, 1, 0]
    # the same pair without the wrap is 0.97 apart
    assert list(ofof.groups(x, 10.0, LL)) == [0, 1, 2]


def test_lowest_index_labels_chain():
    # 3 -- 0 -- 2 chain (1 isolated): the group is labelled by index 0
    x = np.array([[0.5, 0.5, 0.5], [0.1, 0.1, 0.1], [0.54, 0.5, 0.5], [0.46, 0.5, 0.5]],
                 dtype=np.float32)
    assert list(ofof.groups(x, BOX, LL)) == [0, 1, 0, 0]


def _bfs_like_reference(x, L, ll):
    """py/fof.cpp:78-96 and 132-161 with the tree search replaced by a scan of
    every particle: grp[i] = i, then grow each unclaimed particle's group."""
    n = len(x)
    grp = list(range(n))
    ll2 = np.float32(ll) * np.float32(ll)
    for i in range(n):
        if grp[i] != i:
            continue
        q = deque([i])
        while q:
            a = q.popleft()
            d2 = ofof.periodic_d2(x[a:a + 1], x, L)[0]
            for j in range(n):
                if grp[j] != grp[a] and d2[j] < ll2:
                    q.append(j)
                    grp[j] = grp[a]
    return np.array(grp, dtype=np.int32)


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_union_find_equals_breadth_first_growth(seed):
    rng = np.random.default_rng(seed)
    # clumps plus background, so groups of many sizes form
    centres = rng.uniform(0, BOX, (6, 3))
    x = np.concatenate([c + 0.03 * rng.standard_normal((25, 3)) for c in centres] +
                       [rng.uniform(0, BOX, (100, 3))])
    x = np.mod(x, BOX).astype(np.float32)
    x = x[rng.permutation(len(x))]
    g = ofof.groups(x, BOX, LL)
    assert np.array_equal(g, _bfs_like_reference(x, BOX, LL))
    assert 1 < len(np.unique(g)) < len(x)


def test_tree_height():
    assert ofof.tree_height(1, 32) == 0
    assert ofof.tree_height(32, 32) == 0
    assert ofof.tree_height(33, 32) == 1
    assert ofof.tree_height(262144, 32) == 13     # 64^3 particles
    assert ofof.tree_height(332800, 32) == 14     # np_alloc of the reference's 64^3 lpt


def test_tree_layout_hand_derived():
    """16 particles, quota 2, cut cuboid (4, 1, 1).  The cuboid is one array
    halved in depth-first order: root x -> (2,1,1); node 1 x -> (1,1,1);
    node 3 y (tie -> y) -> (1,.5,1); node 4 z -> (1,.5,.5); node 2 x ->
    (.5,.5,.5); node 5 y -> (.5,.25,.5); node 6 z."""
    rng = np.random.default_rng(5)
    n = 16
    x = np.stack([(rng.permutation(n) + rng.uniform(0.1, 0.9, n)) / n for _ in range(3)], 1)
    x = x.astype(np.float32)
    t = ofof.tree_layout(x, 2, boxsize3=(4.0, 1.0, 1.0))
    cut = {0: 0, 1: 0, 3: 1, 4: 2, 2: 0, 5: 1, 6: 2}
    assert sorted(t) == list(range(15))
    for i in range(1, 15):
        parent = (i - 1) // 2
        assert t[i][0] == cut[parent], i
    assert t[0][0] == 0
    ranges = {0: (0, 16), 1: (0, 8), 2: (8, 16), 3: (0, 4), 4: (4, 8), 7: (0, 2), 14: (14, 16)}
    for i, r in ranges.items():
        assert t[i][3:5] == r
    # children split their parent's members at the median of the parent's cut
    for p in range(7):
        k = cut[p]
        lo, hi = t[2 * p + 1][5], t[2 * p + 2][5]
        assert lo | hi == t[p][5] and not lo & hi
        assert max(x[list(lo), k]) < min(x[list(hi), k])
    # extents along the stored axis
    for i, (k, left, right, ib, ie, m) in t.items():
        assert left == float(x[list(m), k].min()) and right == float(x[list(m), k].max())
