"""GPU parity of the HIP engine against the FP64 oracle (called through the C ABI)."""
import numpy as np
import pytest

import decks
import gpu_util as U
import oracle_ctypes as O
import tild

pytestmark = pytest.mark.gpu


def _pair(text, d, seed=7):
    s = tild.System(text, d)
    return s, tild.Engine(s, device=0, seed=seed), O.System(text, d, seed=seed)


CASES = {
    "3d_gauss": lambda t: decks.melt3d(t),
    "3d_mixed": lambda t: decks.melt3d(t, styles=decks.MIXED),
    "3d_odd_grid": lambda t: decks.melt3d(t, grid=(15, 14, 21), box=(6.0, 5.6, 7.0)),
    "2d_gauss": lambda t: decks.melt2d(t),
    "charged": lambda t: decks.charged(t),
    "p0": lambda t: decks.melt3d(t, pmeorder=0),
    "p2": lambda t: decks.melt3d(t, pmeorder=2),
    "p3": lambda t: decks.melt3d(t, pmeorder=3),
    "p4": lambda t: decks.melt3d(t, pmeorder=4),
    "3d_fieldphase": lambda t: _with_fields(decks.melt3d(t), ["fieldphase 1 2 0.5 0 2 1", "fieldphase 1 2 0.3 2 0 1",
                                                              "fieldphase 2 1 0.2 3 0 2"]),
    "2d_fieldphase": lambda t: _with_fields(decks.melt2d(t), ["fieldphase 1 2 0.5 0 1 2", "fieldphase 1 2 0.4 2 2 1"]),
}


def _with_fields(deck, lines):
    text, d = deck
    return text + f"n_fieldphases {len(lines)}\n" + "\n".join(lines) + "\n", d


@pytest.mark.parametrize("case", list(CASES))
def test_density_and_forces(case, tmp_path):
    text, d = CASES[case](tmp_path)
    s, e, o = _pair(text, d)
    print(f"[{case}] ns={s.ns} grid={s.grid_shape}", flush=True)
    f = e.forces()
    fo = o.forces()
    for t in range(e.nfields):
        rho = e.density(t).astype(np.float64)
        ro = o.density(t) if t < s.ntypes else o.charge_density()
        err = np.abs(rho - ro).max() / max(np.abs(ro).max(), 1e-30)
        print(f"  rho[{t}] max rel err {err:.2e}")
        assert err < 2e-6
    ef = U.rel_norm(f, fo)
    print(f"  |df|/|f| = {ef:.2e}  max|df|/rms(f) = {np.abs(f - fo).max() / np.sqrt((fo ** 2).mean()):.2e}")
    assert ef < 1e-5
    e.close()


@pytest.mark.parametrize("case", ["3d_gauss", "3d_mixed", "2d_gauss", "charged"])
def test_properties(case, tmp_path):
    text, d = CASES[case](tmp_path)
    s, e, o = _pair(text, d)
    e.forces()
    p = e.properties()
    o.forces()
    po = o.properties()
    print(f"[{case}] engine {p}\n  oracle {po}")
    for k in ("Upe", "Ubond", "Uelec"):
        assert abs(p[k] - po[k]) <= 1e-5 * max(abs(po[k]), 1e-12), k
    np.testing.assert_allclose(p["U_styles"], po["U_styles"], rtol=1e-5, atol=1e-9)
    scale = np.abs(po["Ptens"]).max()
    assert np.abs(p["Ptens"] - po["Ptens"]).max() <= 1e-5 * scale
    e.close()


def test_coacervate_step0_matches_reference_log(tmp_path):
    text, d = U.coacervate_dir(tmp_path)
    s = tild.System(text, d)
    e = tild.Engine(s, device=0, seed=1)
    e.forces()
    p = e.properties()
    g = U.coacervate_golden()
    print("engine", p, "\nreference data.dat", g)
    assert abs(p["Upe"] / g["Upe"] - 1) < 3e-5
    assert abs(p["Ubond"] / g["Ubond"] - 1) < 3e-5
    assert abs(p["U_styles"][0] / g["Ugauss0"] - 1) < 3e-5
    np.testing.assert_allclose(p["Ptens"], g["Ptens"], rtol=3e-5, atol=3e-7)
    e.close()


@pytest.mark.parametrize("integ", ["EM", "GJF"])
def test_trajectory_20_steps(integ, tmp_path):
    text, d = decks.melt3d(tmp_path, integrator=integ, delt=0.005 if integ == "EM" else 0.01)
    s, e, o = _pair(text, d, seed=21)
    e.step(20)
    o.run(20)
    x = e.positions().astype(np.float64)
    xo = o.positions()
    L = [s.info("L0"), s.info("L1"), s.info("L2")][:s.dim]
    dx = U.min_image(x - xo, L)
    err = np.abs(dx).max()
    print(f"[{integ}] 20-step max |dx| = {err:.2e}")
    assert err < 1e-4
    e.close()


def test_run_twice_bitwise_identical(tmp_path):
    text, d = decks.melt3d(tmp_path, integrator="GJF", delt=0.01)
    runs = []
    for _ in range(2):
        s = tild.System(text, d)
        e = tild.Engine(s, device=0, seed=5)
        e.step(10)
        runs.append((e.positions(), e.density(0)))
        e.close()
    assert np.array_equal(runs[0][0], runs[1][0])
    assert np.array_equal(runs[0][1], runs[1][1])


@pytest.mark.parametrize("P", [1, 2, 3])
def test_bins_bit_exact(P, tmp_path):
    text, d = decks.melt3d(tmp_path, grid=(15, 14, 21), box=(6.0, 5.6, 7.0), pmeorder=P)
    s, e, o = _pair(text, d)
    x = s.positions()
    L = np.array([s.info("L0"), s.info("L1"), s.info("L2")], dtype=np.float32)
    dx = np.array([s.info("dx0"), s.info("dx1"), s.info("dx2")], dtype=np.float32)
    rng = np.random.default_rng(P)
    x[0] = 0.0
    x[1] = L
    k = rng.integers(0, 14, size=(60, 3)).astype(np.float32)
    x[2:62] = k * dx                                  # on grid nodes
    x[62:122] = (k + np.float32(0.5)) * dx            # on node midpoints
    x[122:182] = np.nextafter(k * dx, np.float32(-1))  # one ulp below nodes
    x[182:190] = -1e-6
    e.set_positions(x)
    o.set_positions(x.astype(np.float64))
    eb = e.bins()
    o.forces()
    inds, _ = o.stencil()
    first = inds[:, 0]
    Nx, Ny = s.Nx[0], s.Nx[1]
    ob = np.stack([first % Nx, (first // Nx) % Ny, first // (Nx * Ny)], axis=1)
    bad = np.nonzero((eb != ob).any(axis=1))[0]
    print(f"P={P}: {len(bad)} mismatching bins of {s.ns}")
    assert len(bad) == 0, (bad[:5], eb[bad[:5]], ob[bad[:5]], x[bad[:5]])
    e.close()


SPREAD_CASES = {
    "3d": (lambda t: decks.melt3d(t), [(2, 2), (3, 4), (6, 6)]),
    "3d_odd": (lambda t: decks.melt3d(t, grid=(15, 14, 21), box=(6.0, 5.6, 7.0)), [(2, 3), (7, 7)]),
    "2d": (lambda t: decks.melt2d(t), [(2, 1), (8, 1)]),
    "charged": (lambda t: decks.charged(t), [(2, 2), (4, 3)]),
    "p3": (lambda t: decks.melt3d(t, pmeorder=3), [(4, 4), (6, 6)]),
}


@pytest.mark.parametrize("case", list(SPREAD_CASES))
def test_tiled_spread_bit_identical_to_atomic(case, tmp_path):
    make, shapes = SPREAD_CASES[case]
    text, d = make(tmp_path)
    s = tild.System(text, d)
    ea = tild.Engine(s, device=0, seed=3, spread_mode=1)
    fa = ea.forces()
    rho_a = [ea.density(t) for t in range(ea.nfields)]
    ea.step(3)
    xa = ea.positions()
    for by, bz in shapes:
        et = tild.Engine(s, device=0, seed=3, spread_mode=2, tile_by=by, tile_bz=bz)
        ft = et.forces()
        for t in range(et.nfields):
            assert np.array_equal(et.density(t), rho_a[t]), (case, by, bz, t)
        # the x-axis R2C pairs rows (y, y+1) in both paths when By and Ny are even,
        # which makes everything downstream bit-identical; otherwise rounding-level
        same_pairs = by % 2 == 0 and s.Nx[1] % 2 == 0
        et.step(3)
        if same_pairs:
            assert np.array_equal(ft, fa), (case, by, bz)
            assert np.array_equal(et.positions(), xa), (case, by, bz)
        else:
            assert U.rel_norm(ft, fa) < 1e-6, (case, by, bz)
            assert np.abs(U.min_image(et.positions() - xa, [s.info("L0"), s.info("L1"), s.info("L2")][:s.dim])).max() < 1e-5
        print(f"[{case}] tiles {by}x{bz}: densities bit-identical; forces |df|/|f| = {U.rel_norm(ft, fa):.1e}")
        et.close()
    ea.close()


REORDER_CASES = {
    "3d_gjf": lambda t: decks.melt3d(t, integrator="GJF", delt=0.01, styles=decks.MIXED),
    "2d_em": lambda t: decks.melt2d(t),
    "charged": lambda t: decks.charged(t),
    "p2_atomic": lambda t: decks.melt3d(t, pmeorder=2),
}


@pytest.mark.parametrize("case", list(REORDER_CASES))
def test_spatial_reorder_is_invisible(case, tmp_path):
    # Slots are permuted every 3 steps; everything a caller sees is in site order and
    # bit-identical to a run that never reorders (integer deposition, per-bead gather
    # and bonds, noise keyed by the original site id).
    text, d = REORDER_CASES[case](tmp_path)
    s = tild.System(text, d)
    mode = 1 if case.endswith("atomic") else 0
    print(f"[{case}] ns={s.ns} reorder every 3 steps vs never", flush=True)
    runs = []
    for every in (3, -1):
        e = tild.Engine(s, device=0, seed=11, spread_mode=mode, reorder_every=every)
        e.step(4)
        e.step(9)
        x = e.positions()
        p = e.properties()
        bins = e.bins()
        e.forces()
        f = e.get_forces()
        runs.append((x, p, bins, f, [e.density(t) for t in range(e.nfields)]))
        e.close()
    (x0, p0, b0, f0, r0), (x1, p1, b1, f1, r1) = runs
    assert np.array_equal(x0, x1)
    assert np.array_equal(b0, b1)
    assert np.array_equal(f0, f1)
    for a, b in zip(r0, r1):
        assert np.array_equal(a, b)
    # FP64 property sums see the beads in another order: equal to rounding
    assert abs(p0["Upe"] - p1["Upe"]) <= 1e-12 * abs(p1["Upe"])
    assert np.allclose(p0["Ptens"], p1["Ptens"], rtol=1e-10, atol=1e-14)


def test_reordered_run_matches_oracle(tmp_path):
    text, d = decks.melt3d(tmp_path, integrator="GJF", delt=0.01)
    s = tild.System(text, d)
    e = tild.Engine(s, device=0, seed=3, reorder_every=2)
    o = O.System(text, d, seed=3)
    e.step(10)
    o.run(10)
    L = np.array([s.info("L0"), s.info("L1"), s.info("L2")])
    dx = U.min_image(e.positions().astype(np.float64) - o.positions(), L)
    err = np.abs(dx).max()
    print(f"  10 GJF steps with reorders every 2: max |dx| = {err:.2e}")
    assert err < 1e-4
    e.close()


@pytest.mark.parametrize("case", ["3d", "charged"])
def test_full_bucket_regions_take_the_overflow_path(case, tmp_path):
    # bucket regions sized with no spare slots and never resized: every bead that
    # changes bucket after the sizing lands in the overflow list, and the densities and
    # trajectory stay those of the global-atomic spread bit for bit
    text, d = (decks.melt3d(tmp_path, integrator="GJF", delt=0.01) if case == "3d"
               else decks.charged(tmp_path))
    s = tild.System(text, d)
    ea = tild.Engine(s, device=0, seed=5, spread_mode=1, reorder_every=-1)
    et = tild.Engine(s, device=0, seed=5, spread_mode=2, reorder_every=-1, tile_slack=-1, tile_by=2, tile_bz=2)
    for n in (1, 3, 12):
        ea.step(n)
        et.step(n)
        ea.forces()
        et.forces()
        for t in range(ea.nfields):
            assert np.array_equal(et.density(t), ea.density(t)), (n, t)
        assert np.array_equal(et.positions(), ea.positions()), n
    print(f"[{case}] zero-slack regions: densities and positions bit-identical to the atomic spread")
    ea.close()
    et.close()


@pytest.mark.parametrize("case", ["gjf", "mixed"])
def test_large_tiles_take_the_wide_deposit(case, tmp_path):
    # four 8 x 8-row tiles of 750 beads each run the 512-thread deposit blocks; the
    # densities and the trajectory stay those of the global-atomic spread
    text, d = decks.melt3d(tmp_path, n_chains=250, N=12, box=(8.0, 8.0, 8.0), grid=(16, 16, 16),
                           integrator="GJF" if case == "gjf" else "EM", delt=0.01 if case == "gjf" else 0.005,
                           styles=decks.MIXED if case == "mixed" else None)
    s = tild.System(text, d)
    assert s.ns == 3000
    ea = tild.Engine(s, device=0, seed=2, spread_mode=1)
    et = tild.Engine(s, device=0, seed=2, spread_mode=2, tile_target=2)
    for n in (1, 4):
        ea.step(n)
        et.step(n)
        ea.forces()
        et.forces()
        for t in range(ea.nfields):
            assert np.array_equal(et.density(t), ea.density(t)), (n, t)
        assert np.array_equal(et.positions(), ea.positions()), n
    print(f"[{case}] ns={s.ns}, few large tiles: bit-identical to the atomic spread")
    ea.close()
    et.close()


@pytest.mark.parametrize("case", ["3d", "2d", "charged"])
def test_next_step_claims_follow_the_positions(case, tmp_path):
    # integrating steps claim the next step's bucket slots inside the particle kernel;
    # force evaluations reuse those claims, new positions and spatial reorders drop
    # them.  Every mix of these keeps the densities and the trajectory those of the
    # global-atomic spread bit for bit.
    text, d = (decks.melt3d(tmp_path, integrator="GJF", delt=0.01) if case == "3d"
               else decks.melt2d(tmp_path) if case == "2d" else decks.charged(tmp_path))
    s = tild.System(text, d)
    ea = tild.Engine(s, device=0, seed=6, spread_mode=1, reorder_every=-1)
    et = tild.Engine(s, device=0, seed=6, spread_mode=2, reorder_every=4)
    rng = np.random.default_rng(6)
    L = np.array([s.info("L0"), s.info("L1"), s.info("L2")][:s.dim], dtype=np.float32)

    def same(tag):
        for t in range(ea.nfields):
            assert np.array_equal(et.density(t), ea.density(t)), (tag, t)
        assert np.array_equal(et.positions(), ea.positions()), tag

    for tag, act in [("step 3", lambda e: e.step(3)), ("forces", lambda e: e.forces()),
                     ("forces again", lambda e: e.forces()), ("step 1", lambda e: e.step(1)),
                     ("step 6 over a reorder", lambda e: e.step(6))]:
        ea_f = act(ea)
        et_f = act(et)
        same(tag)
        if tag.startswith("forces"):
            assert np.array_equal(et_f, ea_f), tag
    x = et.positions()
    x = np.mod(x + rng.normal(0.0, 0.3, x.shape).astype(np.float32), L).astype(np.float32)
    x = np.where(x >= L, np.float32(0.0), x).astype(np.float32)
    ea.set_positions(x)
    et.set_positions(x)
    ea.step(2)
    et.step(2)
    ea.forces()
    et.forces()
    same("new positions, step 2")
    print(f"[{case}] fused next-step claims: bit-identical to the atomic spread through "
          "forces, reorders and new positions")
    ea.close()
    et.close()


@pytest.mark.parametrize("case", ["3d_mixed", "charged", "3d_fieldphase"])
def test_plane_forward_matches_separate_passes(case, tmp_path):
    # the forward plane pass pairs rows (y, y+1) in its x transform like an even-height
    # tile deposit, so its half spectra -- and everything downstream -- are bit-identical
    # to the deposit's x transform followed by the separate y pass
    text, d = CASES[case](tmp_path)
    s = tild.System(text, d)
    out = []
    for plane in (0, -1):
        e = tild.Engine(s, device=0, seed=4, fft_plane=plane)
        f = e.forces()
        out.append((f, [e.grid_field(t, k) for t in range(e.nfields) for k in range(s.dim)], e.density(0)))
        e.close()
    (f0, g0, r0), (f1, g1, r1) = out
    assert np.array_equal(r0, r1)
    for a, b in zip(g0, g1):
        assert np.array_equal(a, b)
    assert np.array_equal(f0, f1)
    print(f"[{case}] forward plane pass vs separate passes: bit-identical")


def test_reordered_runs_are_bitwise_reproducible(tmp_path):
    # the tiled reorder sorts each tile's beads by site id, so the slot order -- and with
    # it every FP64 property sum -- repeats exactly from run to run
    text, d = decks.melt3d(tmp_path, integrator="GJF", delt=0.01, styles=decks.MIXED)
    runs = []
    for _ in range(2):
        s = tild.System(text, d)
        e = tild.Engine(s, device=0, seed=8, reorder_every=2)
        e.step(9)
        runs.append((e.positions(), e.properties()))
        e.close()
    assert np.array_equal(runs[0][0], runs[1][0])
    for k in ("Upe", "Ubond"):
        assert runs[0][1][k] == runs[1][1][k]
    assert np.array_equal(runs[0][1]["Ptens"], runs[1][1]["Ptens"])


def test_profiled_steps_are_ordinary_steps(tmp_path):
    # the step profile launches the same kernels with timing events bound to each launch:
    # its steps land where ordinary steps do, bit for bit, and every kernel phase of the
    # tiled path reports a duration
    text, d = decks.melt3d(tmp_path, integrator="GJF", delt=0.01)
    s = tild.System(text, d)
    ea = tild.Engine(s, device=0, seed=9, reorder_every=3)
    eb = tild.Engine(s, device=0, seed=9, reorder_every=3)
    ea.step(2)
    eb.step(2)
    prof = ea.profile(7)
    eb.step(7)
    assert list(prof) == tild.Engine.PHASES and len(prof) == 9
    for k in ("deposit", "last_fused", "x_c2r", "particles", "advance"):
        assert prof[k] > 0.0, (k, prof)
    assert prof["claim"] > 0.0 and prof["reorder"] > 0.0   # reorders at steps 3, 6, 9
    assert np.array_equal(ea.positions(), eb.positions())
    ea.forces()
    eb.forces()
    assert np.array_equal(ea.get_forces(), eb.get_forces())
    print("profile: " + ", ".join(f"{k} {v * 1e3:.1f} us" for k, v in prof.items()))
    ea.close()
    eb.close()


def test_ramped_styles_match_oracle(tmp_path):
    # Gaussian, erf and gaussian-erf prefactors ramped by the reference's update
    # (pair_style.cu:308-328): the engine weights each ramped style's table row per
    # step on the device, so runs of steps replayed from captured graphs (offsets
    # 0..59 in a run) see the same schedule as single steps.  Trajectory and style
    # energies after each run follow the oracle, which rescales its tables after every
    # step as the reference does.
    styles = [("gaussian", 1, 1, 1.0, 0.5), ("gaussian", 2, 2, 1.0, 0.5), ("gaussian", 1, 2, 2.0, 0.5),
              ("erf", 1, 2, 0.1, 1.0, 0.5), ("gaussian_erf", 2, 2, 0.5, 0.6, 0.5)]
    text, d = decks.melt3d(tmp_path, integrator="GJF", delt=0.01, styles=styles)
    for old, new in [("gaussian 1 2 2.0 0.5", "gaussian 1 2 2.0 0.5 ramp 4.0"),
                     ("erf 1 2 0.1 1.0 0.5", "erf 1 2 0.1 1.0 0.5 ramp 0.3"),
                     ("gaussian_erf 2 2 0.5 0.6 0.5", "gaussian_erf 2 2 0.5 0.6 0.5 ramp 0.2"),
                     ("max_steps 1000", "max_steps 200")]:
        assert old in text, old
        text = text.replace(old, new)
    s, e, o = _pair(text, d, seed=13)
    L = [s.info("L0"), s.info("L1"), s.info("L2")][:s.dim]
    done = 0
    for n in (1, 2, 12):
        e.step(n)
        o.run(n)
        done += n
        err = np.abs(U.min_image(e.positions().astype(np.float64) - o.positions(), L)).max()
        p, po = e.properties(), o.properties()
        print(f"  after {done} steps: max |dx| = {err:.2e}, U_styles engine {p['U_styles']} oracle {po['U_styles']}")
        assert err < 1e-4, done
        np.testing.assert_allclose(p["U_styles"], po["U_styles"], rtol=2e-5, atol=1e-9)
        assert abs(p["Upe"] - po["Upe"]) <= 2e-5 * abs(po["Upe"])
    e.close()
    # a 60-step graph replay (offsets 0..59) equals the same steps launched one by one
    g = tild.Engine(s, device=0, seed=13)
    b = tild.Batch(device=0)
    r = b.add(s, 13)
    g.step(61)                       # 60 replayed + 1 eager
    for _ in range(61):
        b.step(1, graph=False)
    assert np.array_equal(g.positions(), b.positions(r))
    np.testing.assert_array_equal(g.properties()["U_styles"], b.properties(r)["U_styles"])
    g.close()
    b.close()


@pytest.mark.parametrize("case", ["gjf", "em_2d"])
def test_restart_from_saved_state_is_exact(case, tmp_path):
    # positions + (step, GJF displacement) + seed restore a run bit for bit: the noise
    # is keyed by (seed, site, step) and the previous step's noise is recomputed; the
    # reference's resume frame (positions only) cannot resume GJF exactly
    text, d = (decks.melt3d(tmp_path, integrator="GJF", delt=0.01, styles=decks.MIXED) if case == "gjf"
               else decks.melt2d(tmp_path))
    s = tild.System(text, d)
    ref = tild.Engine(s, device=0, seed=17, reorder_every=4)
    ref.step(7)
    x7 = ref.positions()
    step, xo = ref.state()
    assert step == 7
    ref.step(13)
    res = tild.Engine(s, device=0, seed=17, reorder_every=3)
    res.step(2)   # a different state and step parity, then overwritten
    res.set_positions(x7)
    res.set_state(step, xo)
    assert res.step_count() == 7
    res.step(13)
    assert np.array_equal(res.positions(), ref.positions())
    # FP64 property sums see the beads in another slot order (the reorders fall on
    # other steps): equal to rounding
    pa, pb = res.properties(), ref.properties()
    for k in ("Upe", "Ubond"):
        assert abs(pa[k] - pb[k]) <= 1e-12 * abs(pb[k]), k
    np.testing.assert_allclose(pa["U_styles"], pb["U_styles"], rtol=1e-12)
    print(f"[{case}] restart at step 7 from saved state: 13 more steps bit-identical")
    ref.close()
    res.close()


GROUP_CASES = {
    # two type groups with different integrators (integrator.cu:24-39, group.cu:44-60)
    "em_and_gjf": ("group A type 1\ngroup B type 2\nintegrator A EM\nintegrator B GJF", False),
    "gjf_and_em_2d": ("group A type 1\ngroup B type 2\nintegrator A GJF\nintegrator B EM", False),
    # only type 2 is integrated: type-1 sites never move
    "one_group": ("group A type 2\nintegrator A GJF", True),
}


@pytest.mark.parametrize("case", list(GROUP_CASES))
def test_group_integrators_match_oracle(case, tmp_path):
    lines, frozen = GROUP_CASES[case]
    if case.endswith("2d"):
        text, d = decks.melt2d(tmp_path, delt=0.01)
    else:
        text, d = decks.melt3d(tmp_path, styles=decks.MIXED, delt=0.01)
    assert "integrator all EM" in text
    text = text.replace("integrator all EM", lines + "\ndiffusivity 2 0.5")
    s, e, o = _pair(text, d, seed=19)
    x0 = s.positions()
    types = s.types()
    L = [s.info("L0"), s.info("L1"), s.info("L2")][:s.dim]
    for n in (1, 9, 10):
        e.step(n)
        o.run(n)
    x = e.positions()
    err = np.abs(U.min_image(x.astype(np.float64) - o.positions(), L)).max()
    print(f"[{case}] groups {int(s.info('n_groups'))}, integrators {int(s.info('n_integrators'))}: "
          f"20-step max |dx| vs FP64 {err:.2e}")
    assert err < 1e-4
    if frozen:
        still = types == 0                    # type 1 (0-based 0) is in no integrator group
        assert np.array_equal(x[still], x0[still])
        assert not np.array_equal(x[~still], x0[~still])
    p, po = e.properties(), o.properties()
    assert abs(p["Upe"] - po["Upe"]) <= 1e-5 * abs(po["Upe"])
    e.close()


def _many_types(tmp, ntypes, grid):
    import tild_gen
    rng = np.random.default_rng(ntypes)
    box = (2.0, 2.0, 0.5 * grid[2])
    pos = rng.random((600, 3)) * np.asarray(box)
    types = np.arange(600) % ntypes + 1
    tild_gen.write_data(tmp / "input.data", pos, types, np.arange(1, 601), [], box, ntypes, 3)
    styles = [("gaussian", t, t, 1.0, 0.5) for t in range(1, ntypes + 1)] + [("gaussian", 1, ntypes, 2.0, 0.5)]
    return tild_gen.deck("input.data", 3, grid, styles).replace("bond 1 harmonic 1.5 0.0\n", ""), tmp


def test_type_count_limit_of_the_fused_pass(tmp_path):
    # the fused last-axis pass holds every field's line in one block: nf * N_last <= 2048
    # (12 types on a 160-point axis fit, 13 do not and are refused)
    (tmp_path / "a").mkdir()
    (tmp_path / "b").mkdir()
    text, d = _many_types(tmp_path / "a", 12, (4, 4, 160))
    s, e, o = _pair(text, d)
    ef = U.rel_norm(e.forces(), o.forces())
    print(f"12 types, N_last 160: |df|/|f| = {ef:.2e}")
    assert ef < 1e-5
    e.close()
    text, d = _many_types(tmp_path / "b", 13, (4, 4, 160))
    s = tild.System(text, d)
    with pytest.raises(tild.TildError) as ei:
        tild.Engine(s, device=0)
    assert ei.value.status == 8
