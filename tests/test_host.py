"""CPU tier: the deck parser and data reader of libtild (host only) against the
oracle's independent transcription, error paths, and the C ABI exports."""
import pathlib
import re
import subprocess

import numpy as np
import pytest

import decks
import gpu_util as U
import oracle_ctypes as O
import tild
import tild_gen

REPO = pathlib.Path(__file__).resolve().parent.parent
EXAMPLES = pathlib.Path("/root/reference/examples")


def _same_system(text, base):
    s = tild.System(text, base)
    o = O.System(text, base, seed=5)
    assert (s.ns, s.dim, s.M, s.ntypes) == (o.ns, o.dim, o.M, o.ntypes)
    assert np.array_equal(s.positions().astype(np.float64), o.positions())
    for a, b in zip(s.bonds(), o.bonds()):
        assert np.array_equal(a, b)
    assert np.array_equal(s.types(), o.types())
    assert np.array_equal(s.molecules(), o.molecules())
    assert np.array_equal(s.charges().astype(np.float64), o.charges())
    for k in ["L0", "L1", "L2", "delt"] + [f"dx{d}" for d in range(s.dim)]:
        assert s.info(k) == o.getd(k), k
    for g in range(int(s.info("n_groups"))):
        assert np.array_equal(s.group_sites(g), o.group_sites(g))
    return s


def test_coacervate_fixture_parses_identically(tmp_path):
    text, d = U.coacervate_dir(tmp_path)
    s = _same_system(text, d)
    assert s.ns == 35588 and int(s.info("n_bonds_total")) == 35154
    assert s.info("do_charges") == 1


@pytest.mark.skipif(not EXAMPLES.exists(), reason="reference examples not mounted")
@pytest.mark.parametrize("ex", ["model-A", "2d-bue", "blend", "Coacervate"])
def test_reference_examples_parse_identically(ex):
    d = EXAMPLES / ex
    _same_system((d / "input").read_text(), d)


@pytest.mark.parametrize("maker", ["melt3d", "melt2d", "charged"])
def test_generated_decks_parse_identically(maker, tmp_path):
    text, d = getattr(decks, maker)(tmp_path)
    _same_system(text, d)


def test_groups_and_integrators(tmp_path):
    text, d = decks.melt3d(tmp_path)
    text = text.replace("integrator all EM", "group A type 1\ngroup B type 2\nintegrator A EM\nintegrator B GJF")
    s = _same_system(text, d)
    assert int(s.info("n_groups")) == 3 and int(s.info("n_integrators")) == 2
    t = s.types()
    assert np.array_equal(s.group_sites(1), np.nonzero(t == 0)[0])


BAD = {
    "unknown keyword": ("Dim 3\nfoo 1\n", tild.TildError, 1),
    "bond before read_data": ("bond 1 harmonic 1 0\n", tild.TildError, 1),
    "pmeorder 5": ("pmeorder 5\n", tild.TildError, 1),
    "missing data file": ("Dim 3\nread_data nothere.data\n", tild.TildError, 1),
    "no read_data": ("Dim 3\nNx 10\n", tild.TildError, 2),
}


@pytest.mark.parametrize("case", list(BAD))
def test_parse_errors_return_status(case):
    text, exc, status = BAD[case]
    with pytest.raises(exc) as ei:
        tild.System(text, "/tmp")
    assert ei.value.status == status


def test_deck_errors_after_read_data(tmp_path):
    text, d = decks.melt3d(tmp_path)
    cases = [
        (text.replace("integrator all EM", "integrator nogroup EM"), 2),
        (text.replace("integrator all EM", "integrator all VV"), 1),
        (text.replace("Nz 12\n", ""), 2),
        (text.replace("n_gaussians 3\ngaussian", "n_gaussians 3\ngaussianx"), 1),
        (text.replace("gaussian 1 2 2.0 0.5", "gaussian 1 2 0.0 0.5 ramp 3.0"), 8),   # ramp from zero
        (text.replace("read_data input.data", "charges 1.0 1.0\nread_data input.data").replace("Dim 3", "Dim 3"), 1),
    ]
    for t, status in cases:
        with pytest.raises(tild.TildError) as ei:
            tild.System(t, d)
        assert ei.value.status == status, t
    # charges after read_data
    with pytest.raises(tild.TildError) as ei:
        tild.System(text + "charges 1.0 1.0\n", d)
    assert ei.value.status == 1


def test_more_than_three_bonds_is_an_error(tmp_path):
    pos = np.random.default_rng(1).random((5, 3)) * 4
    tild_gen.write_data(tmp_path / "input.data", pos, np.ones(5, int), np.ones(5, int),
                        [(1, 2), (1, 3), (1, 4), (1, 5)], (4.0, 4.0, 4.0), 1, 3)
    text = tild_gen.deck("input.data", 3, (8, 8, 8), [("gaussian", 1, 1, 1.0, 0.5)])
    with pytest.raises(tild.TildError) as ei:
        tild.System(text, tmp_path)
    assert ei.value.status == 1


def test_library_exports_every_declared_entry_point():
    header = (REPO / "include" / "tild.h").read_text()
    declared = set(re.findall(r"\b(tild_[a-z0-9_]+)\s*\(", header))
    lib = tild.lib()
    missing = [n for n in sorted(declared) if not hasattr(lib, n)]
    assert not missing, missing
    assert declared == set(tild.EXPORTS)
    out = subprocess.run(["nm", "-D", "--defined-only", str(tild.LIB_PATH)], capture_output=True, text=True).stdout
    exported = set(re.findall(r"\bT (tild_\w+)", out))
    assert declared <= exported


def test_bench_refuses_more_than_one_gpu():
    r = subprocess.run(["python", str(REPO / "bench.py"), "--gpus", "2"], capture_output=True, text=True)
    assert r.returncode != 0 and "one GPU" in r.stderr


def test_tild_run_reports_bad_input(tmp_path):
    exe = pathlib.Path(tild.__file__).resolve().parent / "bin" / "tild-run"
    (tmp_path / "input").write_text("Dim 3\nNx 10\nfrobnicate 3\n")
    r = subprocess.run([str(exe), "-in", "input"], cwd=tmp_path, capture_output=True, text=True, timeout=60)
    assert r.returncode == 1 and "Invalid keyword: frobnicate" in r.stderr
    r = subprocess.run([str(exe)], cwd=tmp_path, capture_output=True, text=True, timeout=60)
    assert r.returncode == 2 and "usage" in r.stderr


def test_fieldphase_decks(tmp_path):
    text, d = decks.melt3d(tmp_path / "a")
    s = tild.System(text + "n_fieldphases 2\nfieldphase 1 2 0.5 0 2 1\nfieldphase 1 2 0.3 3 0 2\n", d)
    assert s.info("n_fieldphases") == 2 and s.info("n_styles") == 3
    bad = [("fieldphase 1 2 0.5 1 0 1", 8),              # BCC: no field in the reference
           ("fieldphase 1 2 0.5 0 3 1", 2),              # lamellae along a 4th axis
           ("fieldphase 1 3 0.5 0 0 1", 1),              # type out of range
           ("fieldphase 1 2 0.5 0 0 1 ramp 1.0", 8)]     # field-phase ramps are outside the engine
    for line, status in bad:
        with pytest.raises(tild.TildError) as ei:
            tild.System(text + "n_fieldphases 1\n" + line + "\n", d)
        assert ei.value.status == status, line
    text2, d2 = decks.melt2d(tmp_path / "b")
    with pytest.raises(tild.TildError) as ei:
        tild.System(text2 + "n_fieldphases 1\nfieldphase 1 2 0.5 3 0 1\n", d2)
    assert ei.value.status == 8
    with pytest.raises(tild.TildError) as ei:   # 2D cylinders must run along the normal
        tild.System(text2 + "n_fieldphases 1\nfieldphase 1 2 0.5 2 0 1\n", d2)
    assert ei.value.status == 2


def test_generator_command_line_writes_runnable_decks(tmp_path):
    # the generator's command line writes an input deck and a data file that both the
    # engine's parser and the oracle read, with the requested sizes
    import subprocess
    import sys
    gen = pathlib.Path(__file__).resolve().parents[1] / "gpu-tild_amd" / "tild_gen.py"
    runs = [
        (["melt", "m3", "--chains", "40", "--N", "10", "--box", "6", "6", "6", "--grid", "12", "12", "12",
          "--integrator", "GJF", "--delt", "0.01", "--chi", "3.0"], 400, 3),
        (["melt", "m2", "--dim", "2", "--chains", "30", "--N", "10", "--box", "10", "8", "--grid", "20", "16"], 300, 2),
        (["coacervate", "c", "--chains-each", "12", "--N", "8", "--counter-pairs", "20", "--total", "600",
          "--box", "7", "7", "8", "--grid", "16", "16", "18"], 600, 3),
    ]
    for args, ns, dim in runs:
        out = tmp_path / args[1]
        subprocess.run([sys.executable, str(gen), args[0], str(out)] + args[2:], check=True,
                       capture_output=True, timeout=60)
        text = (out / "input").read_text()
        s = tild.System(text, out)
        assert s.ns == ns and s.dim == dim, args
        o = O.System(text, out, seed=1)
        np.testing.assert_array_equal(o.positions().astype(np.float32), s.positions().astype(np.float32))
    assert "integrator all GJF" in (tmp_path / "m3" / "input").read_text()
    assert "gaussian 1 2 3.0 0.5" in (tmp_path / "m3" / "input").read_text()
