"""Batched replicas: results equal each replica run alone; graph replays equal eager steps."""
import numpy as np
import pytest

import decks
import tild
import tild_gen

pytestmark = pytest.mark.gpu


def _decks(tmp):
    out = []
    out.append(decks.melt3d(tmp / "a", n_chains=40, N=10, box=(6.0, 6.0, 6.0), grid=(12, 12, 12)))
    out.append(decks.melt3d(tmp / "b", n_chains=30, N=20, box=(7.5, 7.5, 9.0), grid=(15, 15, 18),
                            integrator="GJF", delt=0.01, styles=decks.MIXED))
    out.append(decks.melt2d(tmp / "c"))
    out.append(decks.charged(tmp / "d"))
    out.append(decks.melt3d(tmp / "e", n_chains=20, N=10, box=(6.0, 5.0, 7.0), grid=(12, 10, 14), pmeorder=2))
    return out


@pytest.mark.parametrize("graph", [False, True])
def test_batch_equals_alone(graph, tmp_path):
    systems = [(tild.System(t, d), d) for t, d in _decks(tmp_path)]
    alone = []
    for i, (s, _) in enumerate(systems):
        e = tild.Engine(s, device=0, seed=100 + i)
        e.step(67)
        alone.append((e.positions(), e.density(0), e.properties()))
        e.close()
    b = tild.Batch(device=0)
    ids = [b.add(s, 100 + i) for i, (s, _) in enumerate(systems)]
    b.step(3, graph=graph)
    b.step(64, graph=graph)   # with graph: one replay of the 60-step graph, then 4 eager steps
    for i, r in enumerate(ids):
        assert b.replica_step(r) == 67
        assert np.array_equal(b.positions(r), alone[i][0]), i
        assert np.array_equal(b.density(r, 0), alone[i][1]), i
        p = b.properties(r)
        assert p["Upe"] == alone[i][2]["Upe"], i
    print(f"{len(ids)} uneven replicas (graph={graph}): positions, densities, energies bit-identical to runs alone")
    b.close()


def test_graph_replays_match_eager_with_interleaved_work(tmp_path):
    systems = [(tild.System(t, d), d) for t, d in _decks(tmp_path)[:3]]
    ref = tild.Batch(device=0)
    gb = tild.Batch(device=0)
    rids = [ref.add(s, 7 + i) for i, (s, _) in enumerate(systems)]
    gids = [gb.add(s, 7 + i) for i, (s, _) in enumerate(systems)]
    # runs of 60 or more steps replay the 60-step graph, the rest run eagerly
    for rep, n in enumerate([2, 130, 2, 61]):
        ref.step(n, graph=False)
        gb.step(n, graph=True)
        # eager work and synchronisation between replays
        for r in gids:
            gb.properties(r)
        for a, c in zip(rids, gids):
            assert np.array_equal(ref.positions(a), gb.positions(c)), (rep, a)
    ref.close()
    gb.close()


def test_batch_subsets_and_removal(tmp_path):
    systems = [(tild.System(t, d), d) for t, d in _decks(tmp_path)[:3]]
    b = tild.Batch(device=0)
    ids = [b.add(s, 3) for s, _ in systems]
    b.step(2, replicas=[ids[0], ids[2]])
    assert [b.replica_step(r) for r in ids] == [2, 0, 2]
    b.remove(ids[1])
    b.step(1)
    assert b.replica_step(ids[0]) == 3 and b.replica_step(ids[2]) == 3
    r3 = b.add(systems[1][0], 3)
    b.step(1)
    assert b.replica_step(r3) == 1
    b.close()


@pytest.mark.parametrize("which", ["tiled", "mixed"])
def test_batched_spatial_reorder_is_invisible(which, tmp_path):
    # replicas due for their spatial reorder at the same step are reordered together
    # (one launch per kernel for the set) when all use the tiled spread, one by one
    # otherwise; either way positions are bit-identical to never reordering and the
    # FP64 energies equal to rounding (they sum the beads in another slot order)
    alld = _decks(tmp_path)
    chosen = [alld[0], alld[1], alld[3]] if which == "tiled" else alld
    systems = [tild.System(t, d) for t, d in chosen]
    out = []
    for every in (3, -1):
        # the tiled set is forced onto the tiled spread, so the set-wide reorder runs
        b = tild.Batch(device=0, reorder_every=every, spread_mode=2 if which == "tiled" else 0)
        ids = [b.add(s, 40 + i) for i, s in enumerate(systems)]
        b.step(4, graph=True)
        b.step(9, graph=True)
        b.step(2, graph=False)
        out.append([(b.positions(r), b.properties(r)) for r in ids])
        b.close()
    for i, ((xa, pa), (xb, pb)) in enumerate(zip(*out)):
        assert np.array_equal(xa, xb), (which, i)
        assert abs(pa["Upe"] - pb["Upe"]) <= 1e-12 * abs(pb["Upe"]), (which, i)
    print(f"[{which}] {len(systems)} replicas reordered every 3 steps: positions bit-identical to never reordering")
