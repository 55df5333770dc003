"""GPU parity of the batched ensemble (fs2_batch_*): every box of a batched,
graph-replayed run equals, bit for bit, the same box run alone through the
fs API (fs.lpt.init + fs.pm.force + fs.cola.kick/drift, the reference's
script/cola.py loop), whatever tile shapes the autotuner picks; graph
replays equal the eager launch sequence."""
import numpy as np
import pytest

from oracle import cola as ocola

pytestmark = pytest.mark.gpu

OM = 0.31


def _announce(msg):
    print(f"[case] {msg}", flush=True)


@pytest.fixture(scope="module")
def ps(fs):
    fs.cosmology.init(OM)
    return fs.PowerSpectrum.eisenstein_hu(OM, 0.049, 0.67, 0.96, 0.81)


def _classes():
    # (nc_lpt, nc_pm, nbox, boxsize); 48^3 / 96^3 exercises the radix-3 FFTs
    return [(8, 16, 3, 16.0), (16, 32, 2, 24.0), (48, 96, 2, 48.0)]


def _single_box(fs, ps, nl, nm, L, seed, a0, a1, nstep):
    p = fs.lpt.init(nl, L, a0, ps, seed, "cola")
    fs.pm.init(nm, nm / nl, L)
    fs.pm.set_paint_variant("sorted")
    for a_vel, a_pos in ocola.schedule(a0, a1, nstep):
        fs.pm.force(p)
        fs.cola.kick(p, a_vel)
        fs.cola.drift(p, a_pos)
    return p


def test_batch_matches_single_box_runs_bitwise(fs, ps):
    from fs2_amd.batch import Ensemble
    _announce("batch vs single-box runs")
    classes = _classes()
    a0, a1, nstep = 0.1, 1.0, 3
    ens = Ensemble(classes)
    seeds = list(range(101, 101 + ens.nbox_total))
    ens.init_lpt(ps, a0, seeds)
    tuned = ens.autotune(reps=1)
    for (tile, ms), (nl, nm, nb, L) in zip(tuned, classes):
        assert all(nm % t == 0 for t in tile) and ms > 0
    ens.set_schedule(a0, a1, nstep)
    ens.run(graph=True)
    ens.synchronize()
    k = 0
    for ci, (nl, nm, nb, L) in enumerate(classes):
        for box in range(nb):
            p = _single_box(fs, ps, nl, nm, L, seeds[k], a0, a1, nstep)
            for name in ("x", "v"):
                got = ens.field(ci, box, name)
                ref = getattr(p, name)
                assert np.array_equal(got, ref), (ci, box, name, np.abs(got - ref).max())
            k += 1


def test_generic_mesh_classes_match_single_box_runs_bitwise(fs, ps):
    """Classes on meshes outside the tuned transforms (20^3, 60^3: the
    runtime-planned mixed-radix passes, tiles that no listed shape fits)
    through the autotuner and a graph replay, box by box against the
    single-box loop."""
    from fs2_amd.batch import Ensemble
    _announce("batch generic mesh sizes vs single-box runs")
    classes = [(10, 20, 2, 20.0), (30, 60, 1, 60.0)]
    a0, a1, nstep = 0.1, 1.0, 3
    ens = Ensemble(classes)
    seeds = list(range(301, 301 + ens.nbox_total))
    ens.init_lpt(ps, a0, seeds)
    for (tile, ms), (nl, nm, nb, L) in zip(ens.autotune(reps=1), classes):
        assert all(nm % t == 0 for t in tile) and 0 < ms < 1e3
    ens.set_schedule(a0, a1, nstep)
    ens.run(graph=True)
    ens.synchronize()
    k = 0
    for ci, (nl, nm, nb, L) in enumerate(classes):
        for box in range(nb):
            p = _single_box(fs, ps, nl, nm, L, seeds[k], a0, a1, nstep)
            for name in ("x", "v"):
                got = ens.field(ci, box, name)
                ref = getattr(p, name)
                assert np.array_equal(got, ref), (ci, box, name, np.abs(got - ref).max())
            k += 1


def test_graph_replays_equal_eager_run(fs, ps):
    """Each replay starts from restored ICs, with eager work and host syncs
    between replays; every replay must reproduce the eager result."""
    from fs2_amd.batch import Ensemble
    _announce("batch graph replays vs eager")
    ens = Ensemble(_classes())
    ens.init_lpt(ps, 0.1, list(range(7, 7 + ens.nbox_total)))
    ens.set_tile(0, (8, 8, 16))
    ens.set_schedule(0.1, 1.0, 4)
    ens.run(graph=False)
    ens.synchronize()
    eager = {(c, b): ens.field(c, b, "x") for c in range(3) for b in range(ens.classes[c][2])}
    for rep in range(3):
        ens.reset()
        ens.run(graph=True)
        ens.synchronize()
        for (c, b), ref in eager.items():
            assert np.array_equal(ens.field(c, b, "x"), ref), (rep, c, b)
        # eager work on the same stream between replays
        ens.reset()
        ens.run(graph=False)
        ens.synchronize()
        assert np.array_equal(ens.field(0, 0, "x"), eager[(0, 0)])


def test_tile_shape_does_not_change_results(fs, ps):
    from fs2_amd.batch import Ensemble
    _announce("batch tile shapes")
    ens = Ensemble([(16, 32, 2, 24.0)])
    ens.init_lpt(ps, 0.1, [5, 6])
    ens.set_schedule(0.1, 1.0, 2)
    out = []
    for tile in [(8, 16, 32), (4, 8, 32), (16, 16, 16), (32, 32, 4)]:
        ens.set_tile(0, tile)
        assert ens.tile(0) == tile
        ens.reset()
        ens.run(graph=False)
        ens.synchronize()
        out.append(ens.field(0, 1, "x"))
    for o in out[1:]:
        assert np.array_equal(o, out[0])


def test_batch_errors(fs, ps):
    from fs2_amd.batch import Ensemble
    _announce("batch errors")
    with pytest.raises(RuntimeError):
        Ensemble([(8, 22, 1, 10.0)])          # 22 = 2 x 11 is not an FFT size
    ens = Ensemble([(8, 16, 2, 16.0)])
    with pytest.raises(RuntimeError):
        ens.run()                             # no schedule
    with pytest.raises(RuntimeError):
        ens.autotune()                        # no initial conditions
    with pytest.raises(ValueError):
        ens.init_lpt(ps, 0.1, [1])            # one seed per box
    with pytest.raises(RuntimeError):
        ens.set_tile(0, (3, 8, 8))
