"""tild-run: the command-line driver writes the reference's log, trajectory and binary
files (main.cu:36-293, config_utils.cu:366-402, io_utils.cu:7-134) from the engine."""
import pathlib
import re
import subprocess

import numpy as np
import pytest

import gpu_util
import tild

pytestmark = pytest.mark.gpu

TILD_RUN = pathlib.Path(tild.__file__).resolve().parent / "bin" / "tild-run"


def _g(v):
    """A value as the reference prints a float through an ostream (6 significant digits)."""
    return format(float(np.float32(v)), "g")


def _deck(tmp):
    text, d = gpu_util.coacervate_dir(tmp)
    text = re.sub(r"max_steps \d+", "max_steps 20", text)
    text = re.sub(r"log_freq \d+", "log_freq 10\ntraj_freq 10\ngrid_freq 20", text)
    text = re.sub(r"binary_freq \d+", "binary_freq 10", text) + "rand_seed 11\n"
    (d / "input").write_text(text)
    return text, d


def test_tild_run_outputs(tmp_path):
    assert TILD_RUN.exists(), "tild-run not built"
    text, d = _deck(tmp_path)
    print(f"tild-run on the Coacervate example, 20 steps, in {d}", flush=True)
    r = subprocess.run([str(TILD_RUN), "-in", "input"], cwd=d, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    assert "Step 20 of 20" in r.stdout and "Total run time" in r.stdout

    rows = (d / "data.dat").read_text().splitlines()
    golden = (gpu_util.GOLDEN / "coacervate" / "data.dat").read_text().splitlines()
    assert rows[0] == golden[0].rstrip("\r")                    # same column header
    got0 = np.array([float(v) for v in rows[1].split()])
    ref0 = np.array([float(v) for v in golden[1].split()])
    assert got0[0] == 0 and np.allclose(got0, ref0, rtol=3e-5, atol=0), (got0, ref0)

    # rows at steps 10 and 20 equal the engine's own properties, printed the same way
    s = tild.System(text, d)
    e = tild.Engine(s, device=0)
    ns, dim = s.ns, s.dim
    e.forces()
    frames = {0: e.positions()}
    dens = {}
    for k, step in enumerate((10, 20)):
        e.step(10)
        p = e.properties()
        want = [str(step), _g(p["Upe"]), _g(p["Ubond"])] + [_g(v) for v in p["Ptens"]] + \
               [_g(v) for v in p["U_styles"]]
        assert rows[2 + k].split() == want, (rows[2 + k], want)
        frames[step] = e.positions()
        dens[step] = np.stack([e.density(t).ravel() for t in range(s.ntypes)])
    assert len(rows) == 4

    # trajectory: frames at 0, 10, 20 and the closing frame
    traj = (d / "traj.lammpstrj").read_text().splitlines()
    starts = [i for i, l in enumerate(traj) if l == "ITEM: TIMESTEP"]
    assert [int(traj[i + 1]) for i in starts] == [0, 10, 20, 20]
    last = traj[starts[-1] + 9: starts[-1] + 9 + ns]
    xs = np.array([[float(v) for v in l.split()[3:3 + dim]] for l in last])
    assert np.allclose(xs, frames[20], atol=1e-6)
    assert traj[starts[0] + 8] == "ITEM: ATOMS id type mol x y z q"

    # binary files: headers, then positions and densities at steps 10 and 20, bit-exact
    pb = (d / "positions.bin").read_bytes()
    hdr = np.frombuffer(pb[:8], dtype=np.int32)
    assert list(hdr) == [ns, dim]
    off = 8 + 12 + 8 * ns + 4 + 4 * ns
    pos = np.frombuffer(pb[off:], dtype=np.float32).reshape(2, ns, dim)
    assert np.array_equal(pos[0], frames[10]) and np.array_equal(pos[1], frames[20])
    gb = (d / "grid_densities.bin").read_bytes()
    off = 4 + 4 * dim + 4 * dim + 4
    g = np.frombuffer(gb[off:], dtype=np.float32).reshape(2, s.ntypes, -1)
    assert np.array_equal(g[0], dens[10]) and np.array_equal(g[1], dens[20])
    # rho<t>.dat at step 0 and 20 (the last write wins)
    rho0 = np.loadtxt(d / "rho0.dat")
    assert rho0.shape == (s.M, dim + 1)
    assert np.allclose(rho0[:, -1], dens[20][0], rtol=1e-7, atol=0)
    e.close()



def _normalised_frame(raw):
    """The first frame of a trajectory, lines whitespace-normalised (the reference's
    shipped example has CRLF line ends and one space before the charge column, where
    config_utils.cu:386-396 writes two)."""
    lines = [b" ".join(l.split()) for l in raw.replace(b"\r\n", b"\n").split(b"\n") if l.strip()]
    starts = [i for i, l in enumerate(lines) if l == b"ITEM: TIMESTEP"]
    return lines[:starts[1]] if len(starts) > 1 else lines


def test_tild_run_matches_the_reference_example_files(tmp_path):
    """The headers of positions.bin and grid_densities.bin and the step-0 trajectory
    frame that the reference wrote for examples/Coacervate (init_binary_output,
    io_utils.cu:85-111; write_lammps_traj, config_utils.cu:366-402)."""
    import gzip
    import hashlib
    import json
    text, d = gpu_util.coacervate_dir(tmp_path)
    (d / "input").write_text(re.sub(r"max_steps \d+", "max_steps 1", text))
    r = subprocess.run([str(TILD_RUN), "-in", "input"], cwd=d, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    gold = gpu_util.GOLDEN / "coacervate"
    # binary_freq 25000 > max_steps: both binary files hold their headers only, as the
    # shipped ones do
    assert (d / "grid_densities.bin").read_bytes() == (gold / "grid_densities.bin").read_bytes()
    with gzip.open(gold / "positions.bin.gz", "rb") as f:
        want = f.read()
    got = (d / "positions.bin").read_bytes()
    assert len(got) == len(want) == 427080 and got == want
    meta = json.loads((gold / "traj_step0.json").read_text())
    frame = _normalised_frame((d / "traj.lammpstrj").read_bytes())
    assert [l.decode() for l in frame[:len(meta["head"])]] == meta["head"]
    assert len(frame) == meta["lines"]
    assert hashlib.sha256(b"\n".join(frame)).hexdigest() == meta["sha256_normalised"]
    print("tild-run: positions.bin and grid_densities.bin headers byte-identical, step-0 frame identical "
          "to the reference's example files", flush=True)
