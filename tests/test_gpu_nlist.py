"""Cell-list neighbour lists on the GPU (nlist.cu:13-228) against the oracle and brute force."""
import numpy as np
import pytest

import decks
import oracle_ctypes as O
import tild

pytestmark = pytest.mark.gpu

NL_CASES = {
    "3d": lambda t: decks.melt3d(t, n_chains=120, N=5, box=(6.0, 7.0, 6.0), grid=(12, 14, 12)),
    "2d": lambda t: decks.melt2d(t),
    # two cells along y: neighbouring cells are visited twice and their sites repeat
    "2d_two_cells_y": lambda t: decks.melt2d(t, n_chains=20, box=(10.0, 4.0), grid=(20, 8)),
}


def _brute(x, L, rc, members):
    x = x.astype(np.float32)
    out = {}
    for i in members:
        dxv = x[i] - x[members]
        dxv = np.where(dxv < -L / 2, dxv + L, np.where(dxv > L / 2, dxv - L, dxv)).astype(np.float32)
        r2 = (dxv * dxv).sum(axis=1, dtype=np.float32)
        out[i] = np.sort(np.asarray(members)[r2 < np.float32(rc * rc)])
    return out


@pytest.mark.parametrize("case", list(NL_CASES))
def test_nlist_matches_oracle(case, tmp_path):
    text, d = NL_CASES[case](tmp_path)
    s = tild.System(text, d)
    e = tild.Engine(s, device=0, seed=1)
    o = O.System(text, d, seed=1)
    rc = 1.5
    cell_of, counts, neigh, cmax, nc = e.nlist(rc)
    ocell, ocounts, oneigh, ocmax, onc = o.nlist(rc)
    print(f"[{case}] ns={s.ns} cells={nc} CELL_MAX={cmax} mean list {counts.mean():.1f}", flush=True)
    assert (cmax, nc) == (ocmax, onc)
    assert np.array_equal(cell_of, ocell)
    assert np.array_equal(counts, ocounts)
    for i in range(s.ns):
        assert np.array_equal(neigh[i, :counts[i]], oneigh[i, :counts[i]]), i
    e.close()


def test_nlist_overflow_is_an_error(tmp_path):
    # two cells per axis in 3D: the repeated cells push lists past CELL_MAX, where the
    # reference writes past the end of its list array
    text, d = decks.melt3d(tmp_path, n_chains=8, N=10, box=(4.0, 4.0, 4.0), grid=(8, 8, 8))
    s = tild.System(text, d)
    e = tild.Engine(s, device=0, seed=1)
    with pytest.raises(tild.TildError) as ei:
        e.nlist(1.5)
    assert ei.value.status == 2
    # with room beyond CELL_MAX the lists come back whole, repeats included
    o = O.System(text, d, seed=1)
    cell_of, counts, neigh, cmax, nc = e.nlist(1.5, cap=8 * 1024)
    ocell, ocounts, oneigh, _, _ = o.nlist(1.5)
    assert counts.max() > cmax
    assert np.array_equal(cell_of, ocell) and np.array_equal(counts, ocounts)
    for i in range(s.ns):
        assert np.array_equal(neigh[i, :counts[i]], oneigh[i, :counts[i]]), i
    e.close()


def test_nlist_group_after_reordered_steps(tmp_path):
    text, d = decks.melt3d(tmp_path, n_chains=30, box=(6.0, 7.0, 6.0), grid=(12, 14, 12),
                           integrator="GJF", delt=0.01)
    s = tild.System(text, d)
    e = tild.Engine(s, device=0, seed=2, reorder_every=2)
    e.step(7)                          # slots are permuted now; lists stay by site id
    types = s.types()
    members = np.nonzero(types == 0)[0]
    cell_of, counts, neigh, cmax, nc = e.nlist(1.2, sites=members)
    x = e.positions()
    L = np.array([s.info("L0"), s.info("L1"), s.info("L2")], dtype=np.float32)
    ref = _brute(x, L, 1.2, members)
    outside = np.setdiff1d(np.arange(s.ns), members)
    assert np.all(cell_of[outside] == -1) and np.all(counts[outside] == 0)
    for i in members:
        assert np.array_equal(neigh[i, :counts[i]], ref[i]), i
    e.close()
