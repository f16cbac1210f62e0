"""avg_sk compute on the device (Compute.cu:83-137) against the FP64 oracle, and the
avg_sk_<id>.dat files of tild-run (write_kspace_data, io_utils.cu:60-83)."""
import pathlib
import re
import subprocess

import numpy as np
import pytest

import decks
import oracle_ctypes as O
import tild

pytestmark = pytest.mark.gpu

TILD_RUN = pathlib.Path(tild.__file__).resolve().parent / "bin" / "tild-run"


@pytest.mark.parametrize("case", ["3d", "2d"])
def test_sk_average_matches_oracle(case, tmp_path):
    text, d = decks.melt3d(tmp_path) if case == "3d" else decks.melt2d(tmp_path)
    s = tild.System(text, d)
    e = tild.Engine(s, device=0, seed=9)
    o = O.System(text, d, seed=9)
    for k in range(4):
        e.step(3)
        o.run(3)
        e.sk_accumulate(0, slot=0)
        e.sk_accumulate(1, slot=1)
        o.sk_accumulate(0)
        o.sk_accumulate(1)
    for t in range(2):
        got, n = e.sk_average(t)
        ref, no = o.sk_average(t)
        assert n == no == 4
        err = float(np.abs(got - ref).max() / np.abs(ref).max())
        print(f"avg_sk {case} type {t}: 4 evaluations over 12 steps, max |dS| / max |S| vs FP64 {err:.2e}", flush=True)
        assert err < 1e-5
    e.close()


def test_tild_run_writes_avg_sk(tmp_path):
    text, d = decks.melt3d(tmp_path)
    text = re.sub(r"max_steps \d+", "max_steps 10", text)
    text = re.sub(r"log_freq \d+", "log_freq 5", text) + "compute avg_sk 2 freq 2 wait 3\n"
    (d / "input").write_text(text)
    r = subprocess.run([str(TILD_RUN), "-in", "input", "-seed", "5"], cwd=d, capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stderr
    rows = np.loadtxt(d / "avg_sk_0.dat")
    s = tild.System(text, d)
    assert rows.shape == (s.M - 1, 3 + 4)
    # the same run by hand: evaluations at steps 4, 6, 8 and 10 (step > wait, step % freq == 0)
    e = tild.Engine(s, device=0, seed=5)
    e.forces()
    for step in range(1, 11):
        e.step(1)
        if step > 3 and step % 2 == 0:
            e.sk_accumulate(1, slot=0)
    sk, n = e.sk_average(0)
    assert n == 4
    flat = sk.ravel()[1:]
    scale = np.abs(flat).max()
    assert np.abs(rows[:, 5] - flat.real).max() < 1e-5 * scale
    assert np.abs(rows[:, 6] - flat.imag).max() < 1e-5 * scale
    assert np.allclose(rows[:, 3], np.abs(flat), rtol=1e-5, atol=1e-5 * scale)
    # wavevector columns: 2 pi n / L folded at N / 2, and |k|
    Nx, L = s.Nx[0], s.info("L0")
    i = np.arange(1, s.M)
    nx = i % Nx
    kx = np.where(nx < Nx / 2, nx, nx - Nx) * 2 * np.pi / L
    assert np.allclose(rows[:, 0], kx, atol=2e-6)
    assert np.allclose(rows[:, 4], np.sqrt((rows[:, :3] ** 2).sum(axis=1)), rtol=1e-5)
    e.close()
