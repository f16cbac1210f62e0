"""The block-geometry autotuner (Batch::set_autotune, tild_opts.autotune): results are
bit-identical to the default geometry's, the pick is written to the tuning cache and
read back by later replicas of the same class."""
import numpy as np
import pytest

import decks
import tild

pytestmark = pytest.mark.gpu


def test_autotuned_engine_is_bit_identical_and_cached(tmp_path, monkeypatch):
    cache = tmp_path / "tune.tsv"
    monkeypatch.setenv("TILD_TUNE_CACHE", str(cache))
    text, d = decks.melt3d(tmp_path, n_chains=150, N=20, box=(10.0, 10.0, 12.0), grid=(20, 20, 24), seed=31)
    s = tild.System(text, d)
    ref = tild.Engine(s, device=0, seed=7)
    ref.step(20)
    x_ref = ref.positions()
    ref.close()
    e = tild.Engine(s, device=0, seed=7, autotune=True)
    g = e.geometry()
    print(f"autotune {s.grid_shape}, {s.ns} beads: {g}", flush=True)
    assert g["tuned"] == 1 and g["from_cache"] == 0 and g["step_ms"] > 0
    e.step(20)
    assert np.array_equal(e.positions(), x_ref)
    e.close()
    lines = cache.read_text().splitlines()
    assert len(lines) == 1 and "x24|" in lines[0]
    e2 = tild.Engine(s, device=0, seed=7, autotune=True)
    g2 = e2.geometry()
    assert g2["tuned"] == 1 and g2["from_cache"] == 1
    assert {k: g2[k] for k in ("tiles", "plane", "last_lines", "c2r_rows")} == \
           {k: g[k] for k in ("tiles", "plane", "last_lines", "c2r_rows")}
    e2.step(20)
    assert np.array_equal(e2.positions(), x_ref)
    e2.close()


def test_autotuned_batch_matches_default_batch(tmp_path, monkeypatch):
    monkeypatch.setenv("TILD_TUNE_CACHE", str(tmp_path / "tune.tsv"))
    shapes = [((6.0, 6.0, 7.0), (12, 12, 14), 30), ((8.0, 7.0, 6.0), (16, 14, 12), 45),
              ((5.0, 5.0, 5.0), (10, 10, 10), 20), ((9.0, 9.0, 9.0), (18, 18, 18), 60)]
    systems = []
    for k, (box, grid, nc) in enumerate(shapes):
        t, dd = decks.melt3d(tmp_path / f"r{k}", n_chains=nc, N=10, box=box, grid=grid, seed=50 + k)
        systems.append(tild.System(t, dd))
    out = []
    for tune in (False, True):
        b = tild.Batch(device=0, autotune=tune)
        ids = [b.add(s, 90 + k) for k, s in enumerate(systems)]
        if tune:
            assert all(b.geometry(r)["tuned"] == 1 for r in ids)
        b.step(10, graph=True)
        b.step(7, graph=False)
        out.append([b.positions(r) for r in ids])
        b.close()
    for k in range(len(systems)):
        assert np.array_equal(out[0][k], out[1][k]), k
