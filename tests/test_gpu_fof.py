"""FoF halo finder through the fs API (fs.fof.find_groups, Particles.fof_group,
fs.KdTree): positions copied from the device, groups and kd-tree against the
oracle, and the reference's reordering of the particles (x v dx1 dx2 id
permuted on the device, force not)."""
import numpy as np
import pytest

from oracle import fof as ofof

pytestmark = pytest.mark.gpu

OM = 0.31


def _announce(msg):
    print(f"[case] {msg}", flush=True)


def _line(x0, dx, n):
    return np.array([[x0[a] + i * dx[a] for a in range(3)] for i in range(n)])


def _state(p):
    return {k: getattr(p, k).copy() for k in ("x", "v", "dx1", "dx2", "force", "id")}


def _perm_from_rows(before, after):
    """Index in `before` of each row of `after` (rows must be distinct)."""
    pos = {r.tobytes(): k for k, r in enumerate(np.ascontiguousarray(before))}
    assert len(pos) == len(before)
    return np.array([pos[r.tobytes()] for r in np.ascontiguousarray(after)])


def _check_reordered(before, after):
    # ids identify particles when they are distinct (2LPT lattice ids); appended
    # particles all have id 0, and then their distinct positions do
    if len(np.unique(before["id"])) == len(before["id"]):
        perm = _perm_from_rows(before["id"][:, None], after["id"][:, None])
    else:
        perm = _perm_from_rows(before["x"], after["x"])
        assert np.array_equal(after["id"], before["id"][perm])
    assert sorted(perm) == list(range(len(perm)))
    for k in ("x", "v", "dx1", "dx2"):
        assert np.array_equal(after[k], before[k][perm]), k
    assert np.array_equal(after["force"], before["force"])   # separate array, not permuted
    return perm


@pytest.mark.parametrize("quota", [16, 2])
def test_line_cases_of_reference_tests(fs, quota):
    """test/test_fof.py:40-74: a line spaced 1.1 ll is 10 groups, 0.99 ll one."""
    _announce(f"fof line cases quota={quota}")
    nc, box = 4, 1.0
    dx = box / nc
    ll = 0.2 * box / nc
    for step, ngroups in ((1.1, 10), (0.99, 1)):
        p = fs.Particles(16, box)   # the test's Particles(nc) cannot hold its 10 particles
        p.append(_line([0.5 * dx] * 3, [step * ll, 0, 0], 10))
        nfof = fs.fof.find_groups(p, ll, quota=quota, compute_nfof=True)
        assert np.count_nonzero(nfof) == ngroups
        assert nfof.sum() == 10
        grp = p.fof_group
        assert np.array_equal(grp, ofof.groups(p.x, box, ll))


@pytest.fixture(scope="module")
def lpt_particles(fs):
    fs.cosmology.init(OM)
    ps = fs.PowerSpectrum.eisenstein_hu(OM, 0.049, 0.67, 0.96, 0.81)
    return ps


@pytest.mark.parametrize("quota,box3", [(32, None), (8, (16.0, 16.0, 16.0)), (5, (1.0, 3.0, 2.0))])
def test_lpt_groups_match_brute_force(fs, lpt_particles, quota, box3):
    """Clustered 2LPT particles at a = 1 (16^3 in a 16 Mpc/h box): every
    group label equals the brute-force one on the reordered positions."""
    _announce(f"fof 2lpt 16^3 quota={quota} box3={box3}")
    L = 16.0
    p = fs.lpt.init(16, L, 1.0, lpt_particles, 3, "2lpt")
    p.periodic_wrapup()
    p.set_field("force", np.random.default_rng(1).standard_normal((len(p), 3)).astype(np.float32))
    before = _state(p)
    ll = 0.2 * L / 16
    nfof = fs.fof.find_groups(p, ll, quota=quota, boxsize3=box3, compute_nfof=True)
    after = _state(p)
    _check_reordered(before, after)
    grp = p.fof_group
    ref = ofof.groups(after["x"], L, ll)
    assert np.array_equal(grp, ref)
    assert np.array_equal(nfof, ofof.nfof(ref))
    assert 1 < np.count_nonzero(nfof) < len(p)   # nontrivial grouping
    assert fs.fof.find_groups(p, ll, quota=quota, boxsize3=box3) is None


def test_kdtree_layout_matches_oracle(fs):
    """Distinct coordinates: node ranges, axes, extents and member sets of the
    tree equal the oracle's layout; the tree is sized from the allocation."""
    _announce("fof kd-tree layout")
    rng = np.random.default_rng(7)
    n, L, quota = 3000, 10.0, 7
    x = np.stack([(rng.permutation(n) + rng.uniform(0.1, 0.9, n)) * (L / n) for _ in range(3)], 1)
    p = fs.Particles(4096, L)
    p.append(x)
    for k, name in enumerate(("v", "dx1", "dx2")):   # per-particle tags
        p.set_field(name, (np.arange(3 * n).reshape(n, 3) + 10000 * k).astype(np.float32))
    before = _state(p)
    fs.fof.find_groups(p, 0.05, quota=quota, boxsize3=(10.0, 5.0, 2.5))
    perm = _check_reordered(before, _state(p))
    tree = fs.KdTree()
    assert tree.height == ofof.tree_height(4096, quota)
    assert len(tree._list) == 2 ** (tree.height + 1) - 1
    lay = ofof.tree_layout(before["x"], quota, boxsize3=(10.0, 5.0, 2.5))
    for i, (k, left, right, ib, ie, members) in lay.items():
        node = tree[i]
        assert node[0] == i
        assert (node[1], node[4], node[5]) == (k, ib, ie), i
        assert (node[2], node[3]) == (left, right), i
        assert set(perm[ib:ie].tolist()) == members, i
    for i in range(len(tree._list)):
        if i not in lay:
            assert tree[i][1:] == (0, 0.0, 0.0, 0, 0)


def test_empty_and_errors(fs):
    _announce("fof empty / errors")
    p = fs.Particles(10, 1.0)
    assert fs.fof.find_groups(p, 0.1, compute_nfof=True) is None
    p.append(np.array([[0.1, 0.2, 0.3]]))
    with pytest.raises(TypeError):
        fs.fof.find_groups(p, 0.1, boxsize3=(1, 1, 1))       # ints, as py_fof.cpp:38
    with pytest.raises(ValueError):
        fs.fof.find_groups(p, 0.1, quota=0)
