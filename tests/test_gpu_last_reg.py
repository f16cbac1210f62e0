"""The register-resident FFT passes -- the forward plane pass (k_plane_reg), the fused
last-axis pass (k_last_reg), the y inverse (k_grad_reg) and the inverse plane pass (y
inverse and x C2R in one, k_plane_inv) -- against the LDS-staged ones.

Both forms run the same transforms (forward plan R1 x R2, inverse plan R2 x R1) and
the same coupling sums; the compiler fuses multiplies into additions differently where
one pass keeps in registers what the other stages through LDS, so the two agree to
rounding (1e-6 relative) rather than bit for bit.  The oracle bounds the register
passes at the usual tolerances, and in a batch each replica keeps the bits it has
alone.  The decks cover the two-stage plans the register passes take (axes of 160, 128,
80, 64, 50, 48, 40, 32, 30, 25 and 24 points), one to four types, charges, ramped
styles and 2D; the y inverse and the plane pass take the same plans for their axes (3D
only; the plane pass on square x-y planes).
"""
import pathlib

import numpy as np
import pytest

import decks
import gpu_util as U
import oracle_ctypes as O
import tild
import tild_gen

pytestmark = pytest.mark.gpu

THREE = [("gaussian", 1, 1, 1.0, 0.5), ("gaussian", 2, 2, 1.0, 0.5), ("gaussian", 3, 3, 0.8, 0.5),
         ("gaussian", 1, 2, 2.0, 0.5), ("gaussian", 1, 3, 1.5, 0.5), ("erf", 2, 3, 0.1, 1.0, 0.5)]


def three_types(tmp, grid, box, n_chains=60, N=10, seed=9):
    """Diblock melt with a third type on every third chain's tail."""
    d = pathlib.Path(tmp)
    d.mkdir(parents=True, exist_ok=True)
    rng = np.random.default_rng(seed)
    pos, types, mols, bonds = tild_gen.diblock(rng, n_chains, N, 0.5, box, 3)
    types = np.where((mols % 3 == 0) & (types == 2), 3, types)
    tild_gen.write_data(d / "input.data", pos, types, mols, bonds, box, 3, 3)
    text = tild_gen.deck("input.data", 3, grid, THREE, "GJF", 0.01, seed)
    (d / "input").write_text(text)
    return text, d


def charged_three(tmp, grid, box, n_chains=40, N=10, n_ions=120, seed=10):
    """Polycations (type 1), polyanions (type 2) and neutral counterion-like beads of type 3
    carrying the compensating charges: three types plus the charge field."""
    d = pathlib.Path(tmp)
    d.mkdir(parents=True, exist_ok=True)
    rng = np.random.default_rng(seed)
    pos, types, mols, bonds = tild_gen.diblock(rng, n_chains, N, 0.5, box, 3)
    types = np.where(mols % 2 == 0, 1, 2)
    q = np.where(types == 1, 1.0, -1.0)
    ions = rng.random((n_ions, 3)) * np.asarray(box)
    qi = np.where(np.arange(n_ions) % 2 == 0, 1.0, -1.0)
    pos = np.concatenate([pos, ions])
    types = np.concatenate([types, np.full(n_ions, 3)])
    q = np.concatenate([q, qi])
    mols = np.concatenate([mols, np.arange(n_chains + 1, n_chains + 1 + n_ions)])
    tild_gen.write_data(d / "input.data", pos, types, mols, bonds, box, 3, 3, charges=q)
    styles = [("gaussian_erf", i, j, 2.0, 0.5, 0.5) for i in range(1, 4) for j in range(i, 4)]
    text = tild_gen.deck("input.data", 3, grid, styles, "EM", 0.005, seed, charges=(2.0, 0.5))
    (d / "input").write_text(text)
    return text, d


def ramped(tmp):
    text, d = decks.melt3d(tmp, grid=(8, 8, 160), box=(4.0, 4.0, 40.0), n_chains=80, styles=decks.MIXED,
                           integrator="GJF", delt=0.01)
    for old, new in [("gaussian 1 2 2.0 0.5", "gaussian 1 2 2.0 0.5 ramp 4.0"),
                     ("erf 1 2 0.1 1.0 0.5", "erf 1 2 0.1 1.0 0.5 ramp 0.3")]:
        assert old in text, old
        text = text.replace(old, new)
    return text, d


CASES = {
    "z160": lambda t: decks.melt3d(t, grid=(8, 8, 160), box=(4.0, 4.0, 40.0), n_chains=80),
    "z80_mixed": lambda t: decks.melt3d(t, grid=(10, 8, 80), box=(5.0, 4.0, 20.0), n_chains=60, styles=decks.MIXED,
                                        integrator="GJF", delt=0.01),
    "z64_three": lambda t: three_types(t, grid=(12, 10, 64), box=(6.0, 5.0, 16.0)),
    "z40_charged": lambda t: charged_three(t, grid=(12, 12, 40), box=(6.0, 6.0, 20.0)),
    "z160_ramp": ramped,
    "2d_y80": lambda t: decks.melt2d(t, grid=(20, 80), box=(10.0, 20.0)),
    "2d_y40": lambda t: decks.melt2d(t, grid=(15, 40), box=(6.0, 10.0)),
    "y80_z40": lambda t: decks.melt3d(t, grid=(10, 80, 40), box=(5.0, 20.0, 10.0), n_chains=80),
    "y64_z64_three": lambda t: three_types(t, grid=(9, 64, 64), box=(4.5, 16.0, 16.0)),
    "y40_z20_4types": lambda t: decks.charged(t, grid=(12, 40, 20)),   # z = 20: staged z pass
    "z64_4types_charged": lambda t: decks.charged(t, grid=(12, 12, 64)),
    "x80_y40_z64": lambda t: decks.melt3d(t, grid=(80, 40, 64), box=(20.0, 10.0, 16.0), n_chains=100),
    "x64_z40_charged": lambda t: charged_three(t, grid=(64, 12, 40), box=(16.0, 6.0, 20.0)),
    "2d_x40_y40": lambda t: decks.melt2d(t, grid=(40, 40), box=(10.0, 10.0)),
    "2d_x160_y16": lambda t: decks.melt2d(t, grid=(160, 16), box=(40.0, 8.0)),
    # one register pass (with the y inverse) at a time
    "plane_x40_y40_z20_4types": lambda t: decks.charged(t, grid=(40, 40, 20)),
    "y40_x80_z20": lambda t: decks.melt3d(t, grid=(80, 40, 20), box=(20.0, 10.0, 5.0), n_chains=80),
    "plane_x64_y64_z10": lambda t: decks.melt3d(t, grid=(64, 64, 10), box=(16.0, 16.0, 5.0), n_chains=100),
    "x48_y48_z30": lambda t: decks.melt3d(t, grid=(48, 48, 30), box=(12.0, 12.0, 7.5), n_chains=100),
    "x24_y32_z25_three": lambda t: three_types(t, grid=(24, 32, 25), box=(6.0, 8.0, 6.25)),
    "2d_x50_y128": lambda t: decks.melt2d(t, grid=(50, 128), box=(12.5, 32.0)),
    # odd square planes: the last row's z derivative has no partner row
    "inverse_odd_x25_y25_z20": lambda t: decks.melt3d(t, grid=(25, 25, 20), box=(12.5, 12.5, 10.0), n_chains=100),
}

# two-stage plans the register passes take, and the variants compiled in
PLANS = {160: (16, 10), 80: (10, 8), 64: (8, 8), 40: (8, 5), 128: (16, 8), 50: (10, 5), 48: (16, 3), 32: (8, 4),
         30: (10, 3), 25: (5, 5), 24: (8, 3)}


def expected_flags(s, g):
    nx = s.grid_shape[::-1]                     # (Nx, Ny[, Nz])
    nf = s.ntypes + (1 if s.info("do_charges") else 0)
    D = s.dim
    last = nx[-1] in PLANS and s.ntypes <= (2 if PLANS[nx[-1]][0] == 16 else 4)
    grad = D == 3 and nx[1] in PLANS
    # the forward plane pass (where the replica runs one: tiled 3D, even Ny and tile
    # height) on square x-y planes
    plane = g["plane"] == 1 and nx[0] == nx[1] and nx[0] in PLANS
    return int(last), int(grad), int(plane)


def expected_inverse_plane(s):
    # the y inverse and x C2R as one pass: square x-y planes of a two-stage plan whose
    # half-plane columns times the larger radix fit the block
    nx = s.grid_shape[::-1]
    if s.dim != 3 or nx[0] != nx[1] or nx[0] not in PLANS:
        return 0
    return int((nx[0] // 2 + 1) * max(PLANS[nx[0]]) <= 512)


@pytest.mark.parametrize("case", list(CASES))
def test_register_pass_equals_staged_pass(case, tmp_path):
    text, d = CASES[case](tmp_path)
    s = tild.System(text, d)
    reg = tild.Engine(s, device=0, seed=5)
    staged = tild.Engine(s, device=0, seed=5, fft_staged=True)
    gr, gs = reg.geometry(), staged.geometry()
    print(f"[{case}] ns={s.ns} grid={s.grid_shape} last_reg={gr['last_reg']} grad_reg={gr['grad_reg']} "
          f"plane_reg={gr['plane_reg']}", flush=True)
    flags = (gr["last_reg"], gr["grad_reg"], gr["plane_reg"])
    assert flags == expected_flags(s, gr), (flags, gr)
    assert gr["plane_inv"] == expected_inverse_plane(s), gr
    print(f"  plane_inv={gr['plane_inv']}")
    if case.startswith("plane"):
        assert gr["plane_reg"] == 1
    assert gs["last_reg"] == 0 and gs["grad_reg"] == 0 and gs["plane_reg"] == 0 and gs["plane_inv"] == 0
    fr, fs = reg.forces(), staged.forces()
    ef = U.rel_norm(fr, fs)
    print(f"  |f_reg - f_staged|/|f| = {ef:.2e}, bit-identical: {np.array_equal(fr, fs)}")
    assert ef < 1e-6
    for e in (reg, staged):
        e.step(7)
    xr, xs = reg.positions(), staged.positions()
    L = [s.info("L0"), s.info("L1"), s.info("L2")][:s.dim]
    assert np.abs(U.min_image(xr.astype(np.float64) - xs, L)).max() < 1e-5
    pr, ps = reg.properties(), staged.properties()
    assert abs(pr["Upe"] - ps["Upe"]) <= 1e-6 * abs(ps["Upe"])
    for e in (reg, staged):
        e.close()


@pytest.mark.parametrize("case", ["z160", "z64_three", "z40_charged", "z64_4types_charged", "2d_y80", "y80_z40",
                                  "y64_z64_three",
                                  "x80_y40_z64", "2d_x40_y40", "2d_x160_y16", "plane_x40_y40_z20_4types",
                                  "plane_x64_y64_z10", "x48_y48_z30", "x24_y32_z25_three", "2d_x50_y128",
                                  "inverse_odd_x25_y25_z20"])
def test_register_pass_matches_oracle(case, tmp_path):
    text, d = CASES[case](tmp_path)
    s = tild.System(text, d)
    e = tild.Engine(s, device=0, seed=7)
    o = O.System(text, d, seed=7)
    g = e.geometry()
    assert (g["last_reg"], g["grad_reg"], g["plane_reg"]) == expected_flags(s, g)
    assert g["plane_inv"] == expected_inverse_plane(s)
    f, fo = e.forces(), o.forces()
    ef = U.rel_norm(f, fo)
    p, po = e.properties(), o.properties()
    print(f"[{case}] |df|/|f| = {ef:.2e}  Upe {p['Upe']:.9e} oracle {po['Upe']:.9e}")
    assert ef < 1e-5
    assert abs(p["Upe"] - po["Upe"]) <= 1e-5 * abs(po["Upe"])
    e.close()


def test_register_pass_in_a_mixed_batch(tmp_path):
    # replicas of either pass share one batch: each runs its own pass, and every
    # replica equals the same deck run alone
    a = decks.melt3d(tmp_path / "a", grid=(8, 8, 160), box=(4.0, 4.0, 40.0), n_chains=80)
    b = decks.melt3d(tmp_path / "b")                       # 12^3: the staged pass
    c = charged_three(tmp_path / "c", grid=(12, 12, 40), box=(6.0, 6.0, 20.0))
    sys_ = [tild.System(t, d) for t, d in (a, b, c)]
    bt = tild.Batch(device=0)
    ids = [bt.add(s, 11 + k) for k, s in enumerate(sys_)]
    assert [bt.geometry(r)["last_reg"] for r in ids] == [1, 0, 1]
    bt.step(9)
    for k, s in enumerate(sys_):
        e = tild.Engine(s, device=0, seed=11 + k)
        e.step(9)
        assert np.array_equal(bt.positions(ids[k]), e.positions()), k
        e.close()
    bt.close()
