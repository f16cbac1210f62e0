"""avg_sk compute (Compute.cu:20-137): deck grammar and the oracle's restatement.

The oracle's average of rho(k) rho(k) / M is pinned by (a) the closed form of one bead
sitting on a grid node with pmeorder 0, whose density is a single node of height
1/gvol, so rho(k) = exp(-i k.r0) / gvol and its square is exp(-2i k.r0) / gvol^2, and
(b) numpy's FFT of the oracle's own densities for a melt, averaged over several
steps.  The GPU tier compares the engine's accumulator with this oracle."""
import pathlib

import numpy as np
import pytest

import decks
import oracle_ctypes as O
import tild
import tild_gen


def test_compute_lines_parse(tmp_path):
    text, d = decks.melt3d(tmp_path)
    s = tild.System(text + "compute avg_sk 1 freq 10 wait 5\ncompute avg_sk 2\n", d)
    assert s.computes() == [(0, 10, 5), (1, 100, 0)]
    assert tild.System(text, d).computes() == []


@pytest.mark.parametrize("line,where", [
    ("compute avg_sk 3", "after"),           # only two types
    ("compute avg_sk 0", "after"),
    ("compute avg_sk 1", "before"),          # the reference checks against the types read so far
    ("compute avg_rho 1", "after"),
])
def test_compute_errors(tmp_path, line, where):
    text, d = decks.melt3d(tmp_path)
    bad = text.replace("read_data input.data", line + "\nread_data input.data") if where == "before" else text + line + "\n"
    with pytest.raises(tild.TildError) as ei:
        tild.System(bad, d)
    assert ei.value.status == 1
    msg = str(ei.value)
    assert ("Undefined compute style!" in msg) if "avg_rho" in line else ("Invalid type to calculate avg_sk" in msg)


def _one_bead(tmp, node, grid=(8, 6, 10), box=(4.0, 3.0, 5.0)):
    d = pathlib.Path(tmp)
    dx = [box[i] / grid[i] for i in range(3)]
    # pmeorder 0 puts the bead's whole weight on bin int((x + dx/2) / dx): a bead at a
    # node position lands on that node
    pos = np.array([[node[i] * dx[i] for i in range(3)]])
    tild_gen.write_data(d / "input.data", pos, np.ones(1, int), np.ones(1, int), [], box, 1, 3)
    text = tild_gen.deck("input.data", 3, grid, [("gaussian", 1, 1, 1.0, 0.5)], pmeorder=0)
    return text, d, dx


def test_oracle_sk_of_one_bead_on_a_node(tmp_path):
    node = (3, 1, 7)
    text, d, dx = _one_bead(tmp_path, node)
    o = O.System(text, d, seed=1)
    o.forces()
    gvol = float(np.prod(np.float32(dx).astype(np.float64)))
    rho = o.density(0)
    assert np.count_nonzero(rho) == 1 and rho[node[2], node[1], node[0]] == pytest.approx(1.0 / gvol, rel=1e-12)
    o.sk_accumulate(0)
    sk, n = o.sk_average(0)
    assert n == 1
    Nz, Ny, Nx = rho.shape
    kz, ky, kx = np.meshgrid(np.arange(Nz), np.arange(Ny), np.arange(Nx), indexing="ij")
    phase = 2 * np.pi * (kx * node[0] / Nx + ky * node[1] / Ny + kz * node[2] / Nz)
    expect = np.exp(-2j * phase) / gvol ** 2 / rho.size
    assert np.abs(sk - expect).max() < 1e-12 * np.abs(expect).max()
    # the square, not |rho(k)|^2: away from k = 0 the phases do not cancel
    assert np.abs(sk.imag).max() > 0.1 * np.abs(expect).max()


def test_oracle_sk_average_matches_numpy_fft(tmp_path):
    text, d = decks.melt3d(tmp_path)
    o = O.System(text, d, seed=2)
    ref = np.zeros((2,) + o.grid_shape, dtype=complex)
    for k in range(3):
        o.run(4)
        for t in range(2):
            o.sk_accumulate(t)
            rk = np.fft.fftn(o.density(t))
            ref[t] += rk * rk / o.M
    for t in range(2):
        sk, n = o.sk_average(t)
        assert n == 3
        want = ref[t] / 3
        assert np.abs(sk - want).max() < 1e-12 * np.abs(want).max()
    empty = O.System(text, d, seed=2)
    sk0, n0 = empty.sk_average(1)
    assert n0 == 0 and not np.any(sk0)
