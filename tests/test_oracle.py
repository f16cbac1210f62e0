"""CPU tier: pin the FP64 oracle against the reference's golden record, known-answer
vectors, closed forms and brute force (no GPU needed)."""
import json
import math
import pathlib

import numpy as np
import pytest

import decks
import gpu_util as U
import oracle_ctypes as O
import tild_gen

GOLD = pathlib.Path(__file__).resolve().parent / "golden"


# ---------------------------------------------------------------- primitives
def test_philox_known_answers():
    kat = json.loads((GOLD / "philox_kat.json").read_text())
    for c in kat["cases"]:
        got = O.philox(c["ctr"], c["key"])
        assert [f"{v:08x}" for v in got] == c["out"]


def test_normals_moments():
    z = np.array([O.normals(7, i, 3) for i in range(40000)]).ravel()
    assert abs(z.mean()) < 0.01
    assert abs(z.var() - 1.0) < 0.02
    # fourth moment of a standard normal is 3
    assert abs((z ** 4).mean() - 3.0) < 0.1
    # independent streams for different steps and sites
    a = np.array([O.normals(7, i, 3)[0] for i in range(5000)])
    b = np.array([O.normals(7, i, 4)[0] for i in range(5000)])
    assert abs(np.corrcoef(a, b)[0, 1]) < 0.05


def test_fft_golden_vectors():
    g = np.load(GOLD / "fft_vectors.npz")
    for k in range(int(g["n"])):
        a, ref = g[f"in_{k}"], g[f"fwd_{k}"]
        got = O.fft(a.astype(np.complex128), -1)
        assert np.abs(got - ref).max() <= 1e-12 * np.abs(ref).max(), a.shape
        back = O.fft(got, +1) / a.size
        assert np.abs(back - a).max() < 1e-12


@pytest.mark.parametrize("n", [2, 3, 4, 5, 6, 7, 8, 9, 11, 12, 13, 16, 25, 27, 49])
def test_fft_naive_dft(n):
    rng = np.random.default_rng(n)
    a = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    k = np.arange(n)
    F = np.exp(-2j * np.pi * np.outer(k, k) / n)
    assert np.abs(O.fft(a, -1) - F @ a).max() < 1e-11 * max(1.0, np.abs(a).sum())


def test_fft_parseval_3d():
    rng = np.random.default_rng(2)
    a = rng.standard_normal((6, 5, 7))
    A = O.fft(a.astype(np.complex128), -1)
    assert abs((np.abs(A) ** 2).sum() / a.size - (a ** 2).sum()) < 1e-10


# ---------------------------------------------------------------- golden record
def test_coacervate_step0_reproduces_reference_log(tmp_path):
    text, d = U.coacervate_dir(tmp_path)
    o = O.System(text, d, seed=1)
    o.forces()
    p = o.properties()
    g = U.coacervate_golden()
    # 6 printed digits and FP32 host sums in the reference: ~3e-5 relative
    assert abs(p["Upe"] / g["Upe"] - 1) < 3e-5
    assert abs(p["Ubond"] / g["Ubond"] - 1) < 3e-5
    assert abs(p["U_styles"][0] / g["Ugauss0"] - 1) < 3e-5
    np.testing.assert_allclose(p["Ptens"], g["Ptens"], rtol=3e-5, atol=3e-7)


# ---------------------------------------------------------------- spread
@pytest.mark.parametrize("P", [0, 1, 2, 3])
def test_spread_sum_rule(P, tmp_path):
    text, d = decks.melt3d(tmp_path, pmeorder=P)
    o = O.System(text, d, seed=1)
    o.forces()
    gv = o.getd("gvol")
    nt = np.bincount(o.types(), minlength=o.ntypes)
    for t in range(o.ntypes):
        assert abs(o.density(t).sum() * gv - nt[t]) < 1e-9 * nt[t]
    _, W = o.stencil()
    np.testing.assert_allclose(W.sum(axis=1) * gv, 1.0, atol=1e-12)


def _cardinal_bspline(k, t):
    """Cox-de Boor cardinal B-spline M_k(t) (order k, support [0, k])."""
    if k == 1:
        return 1.0 if 0.0 <= t < 1.0 else 0.0
    return (t * _cardinal_bspline(k - 1, t) + (k - t) * _cardinal_bspline(k - 1, t - 1.0)) / (k - 1)


@pytest.mark.parametrize("P", [0, 1, 2, 3, 4])
def test_spread_weights_are_cardinal_bsplines(P, tmp_path):
    # stencil node j of a coordinate sits (j - P/2 - sx) grid units from the particle for
    # even P (sx = x/dx - g) and (j - (P-1)/2 - 1/2 - sx) for odd P (sx = x/dx - g - 1/2),
    # so in both cases its weight is M_{P+1}(1/2 + j - sx).  The reference's order-4 weights use
    # sx3 = sx2*sx2 (device_grid_utils.cu:39) and must differ from the spline there.
    text, d = decks.melt3d(tmp_path, n_chains=10, pmeorder=P)
    o = O.System(text, d, seed=1)
    o.forces()
    _, W = o.stencil()
    gv = o.getd("gvol")
    x = o.positions().astype(np.float32)
    dxs = [np.float32(o.getd(f"dx{a}")) for a in range(3)]
    worst = 0.0
    for i in range(o.ns):
        w1 = []
        for a in range(3):
            xf, dxf = x[i, a], dxs[a]
            if P % 2 == 0:
                g = int(np.float32(xf + np.float32(0.5) * dxf) / dxf)
                sx = (float(xf) - g * float(dxf)) / float(dxf)
            else:
                g = int(np.float32(xf / dxf))
                sx = (float(xf) - (g + 0.5) * float(dxf)) / float(dxf)
            w1.append([_cardinal_bspline(P + 1, 0.5 + j - sx) for j in range(P + 1)])
        ref = np.einsum("i,j,k->ijk", *w1).ravel()
        worst = max(worst, np.abs(W[i] * gv - ref).max())
    if P <= 3:
        assert worst < 1e-12
    else:
        assert worst > 1e-3
    np.testing.assert_allclose(W.sum(axis=1) * gv, 1.0, atol=1e-12)


def test_cic_particle_on_node_is_exact(tmp_path):
    text, d = decks.melt3d(tmp_path)
    o = O.System(text, d, seed=1)
    x = o.positions()
    dx = o.getd("dx0")
    x[0] = [2 * dx, 3 * dx, 5 * dx]
    o.set_positions(x)
    o.forces()
    inds, W = o.stencil()
    gv = o.getd("gvol")
    assert abs(W[0].max() * gv - 1.0) < 1e-12 and abs(W[0].sum() * gv - 1.0) < 1e-12


# ---------------------------------------------------------------- forces
def _sys_without_bonds(tmp, styles, n=2, box=(20.0, 20.0, 20.0), grid=(100, 100, 100), ntype=(1, 2)):
    d = pathlib.Path(tmp)
    d.mkdir(parents=True, exist_ok=True)
    # off-node positions: a coordinate whose FP32 x/dx rounds onto a node gets a weight
    # of about -1e-7, and the rho > 0 mask of the gather (device_grid_utils.cu:640) then
    # drops that node on one side only, breaking Newton's third law at ~1e-6
    pos = np.array([[10.05, 10.13, 9.93], [11.27, 10.41, 9.66]])[:n]
    tild_gen.write_data(d / "input.data", pos, np.array(ntype[:n]), np.arange(1, n + 1), [], box, 2, 3)
    text = tild_gen.deck("input.data", 3, grid, styles)
    return text.replace("bond 1 harmonic 1.5 0.0\n", ""), d


def test_momentum_conservation(tmp_path):
    text, d = decks.melt3d(tmp_path, styles=decks.MIXED)
    o = O.System(text, d, seed=1)
    f = o.forces()
    assert np.abs(f.sum(axis=0)).max() < 1e-9 * np.abs(f).sum()
    text2, d2 = decks.charged(tmp_path / "c")
    o2 = O.System(text2, d2, seed=1)
    f2 = o2.forces()
    assert np.abs(f2.sum(axis=0)).max() < 1e-9 * np.abs(f2).sum()


def test_two_particle_gaussian_force_matches_analytic(tmp_path):
    # u(r) = A (2 pi s^2)^{-3/2} exp(-r^2 / 2 s^2) for u(k) = A exp(-k^2 s^2 / 2); the grid
    # force (CIC + ik differentiation) approaches -grad u at second order in dx.
    A, sigma = 2.0, 1.0
    errs = []
    for n in (40, 80):
        text, d = _sys_without_bonds(tmp_path / str(n), [("gaussian", 1, 2, A, sigma)], grid=(n, n, n))
        o = O.System(text, d, seed=1)
        f = o.forces()
        x = o.positions()
        r = x[0] - x[1]
        rr = np.linalg.norm(r)
        u = A / (2 * math.pi * sigma ** 2) ** 1.5 * math.exp(-rr ** 2 / (2 * sigma ** 2))
        fa = u * r / sigma ** 2
        errs.append(np.linalg.norm(f[0] - fa) / np.linalg.norm(fa))
        assert np.linalg.norm(f[0] + f[1]) < 1e-9 * np.linalg.norm(fa)
    assert errs[1] < 0.05
    assert errs[0] / errs[1] > 3.0, errs     # halving dx cuts the error ~4x


def _radial_gauss_conv(f, r, sigma, smax):
    """(G_sigma * f)(r) for an isotropic f supported on [0, smax] (3D radial convolution)."""
    from scipy import integrate

    def g(s):
        return s * f(s) * (math.exp(-(r - s) ** 2 / (2 * sigma ** 2)) - math.exp(-(r + s) ** 2 / (2 * sigma ** 2)))
    v, _ = integrate.quad(g, 0.0, smax, limit=200)
    return v / (r * sigma * math.sqrt(2 * math.pi))


@pytest.mark.parametrize("kind", ["gaussian_erf", "erf"])
def test_two_particle_erf_energies_match_real_space(kind, tmp_path):
    # gaussian_erf: u(k) = A exp(-k^2 s^2/2) h(k), erf: A exp(-k^2 s^2/2) h(k)^2 with h the
    # transform of a sphere of radius Rp (pair_style_gaussian_erf.cu:82-115,
    # pair_style_erf.cu:122-157).  In real space that is A times a Gaussian smoothing
    # of the sphere indicator, resp. of the two-sphere overlap volume
    # pi (4R + d)(2R - d)^2 / 12, evaluated here by radial quadrature.
    A, R, sig = 1.5, 1.0, 0.7
    if kind == "erf":
        f, smax = (lambda s: math.pi / 12 * (4 * R + s) * (2 * R - s) ** 2 if s < 2 * R else 0.0), 2 * R
    else:
        f, smax = (lambda s: 1.0 if s < R else 0.0), R
    errs = []
    for n in (40, 80):
        text, d = _sys_without_bonds(tmp_path / str(n), [(kind, 1, 2, A, R, sig)], grid=(n, n, n))
        o = O.System(text, d, seed=1)
        o.forces()
        x = o.positions()
        ua = A * _radial_gauss_conv(f, float(np.linalg.norm(x[0] - x[1])), sig, smax)
        errs.append(abs(o.properties()["U_styles"][0] / ua - 1.0))
    assert errs[1] < 0.03 and errs[0] / errs[1] > 2.0, errs


def test_two_charge_force_matches_smeared_coulomb(tmp_path):
    # The force path smears twice (device_array_utils.cu:213, 237-238): the field is that
    # of 4 pi lB / k^2 exp(-k^2/a^2), i.e. lB erf(a r / 2) / r in real space.
    lB, a = 1.0, 2.0
    errs = []
    for n in (40, 80):
        dd = tmp_path / str(n)
        dd.mkdir()
        pos = np.array([[10.05, 10.13, 9.93], [11.27, 10.41, 9.66]])
        tild_gen.write_data(dd / "input.data", pos, np.array([1, 2]), np.arange(1, 3), [], (20.0, 20.0, 20.0),
                            2, 3, charges=np.array([1.0, -1.0]))
        text = tild_gen.deck("input.data", 3, (n, n, n), [("gaussian", 1, 2, 0.0, 1.0)], charges=(lB, a))
        o = O.System(text.replace("bond 1 harmonic 1.5 0.0\n", ""), dd, seed=1)
        f = o.forces()
        x = o.positions()
        rv = x[0] - x[1]
        r = float(np.linalg.norm(rv))
        b = a / 2
        dV = -lB * (2 * b / math.sqrt(math.pi) * math.exp(-b * b * r * r) / r - math.erf(b * r) / r ** 2)
        fa = -dV * rv / r
        errs.append(np.linalg.norm(f[0] - fa) / np.linalg.norm(fa))
        assert np.linalg.norm(f[0] + f[1]) < 1e-9 * np.linalg.norm(fa)
    assert errs[1] < 0.05 and errs[0] / errs[1] > 2.5, errs


def test_force_is_minus_energy_gradient(tmp_path):
    # ik-differentiated forces are not the exact gradient of the gridded energy; the
    # mismatch vanishes as dx -> 0 (checked at two resolutions)
    styles = [("gaussian", 1, 2, 1.0, 1.0), ("erf", 1, 2, 0.3, 1.0, 1.0)]
    errs = []
    for n in (40, 80):
        text, d = _sys_without_bonds(tmp_path / str(n), styles, grid=(n, n, n))
        o = O.System(text, d, seed=1)
        f = o.forces()
        x0 = o.positions()
        h = 1e-3
        g = np.zeros(3)
        for ax in range(3):
            e = []
            for s in (+1, -1):
                x = x0.copy()
                x[0, ax] += s * h
                o.set_positions(x)
                o.forces()
                e.append(o.properties()["Upe"])
            g[ax] = (e[0] - e[1]) / (2 * h)
        errs.append(np.linalg.norm(-g - f[0]) / np.linalg.norm(f[0]))
    assert errs[1] < 0.05 and errs[0] / errs[1] > 2.5, errs


def test_bond_force_is_energy_derivative(tmp_path):
    d = pathlib.Path(tmp_path)
    rng = np.random.default_rng(3)
    pos = 5 + rng.normal(0, 0.6, size=(6, 3))
    tild_gen.write_data(d / "input.data", pos, np.ones(6, int), np.ones(6, int),
                        [(1, 2), (2, 3), (3, 4), (2, 5), (5, 6)], (10.0, 10.0, 10.0), 1, 3)
    text = tild_gen.deck("input.data", 3, (10, 10, 10), [], bond_k=1.7, bond_r0=0.4)
    o = O.System(text, d, seed=1)
    f = o.forces()
    x0 = o.positions()
    h = 1e-5
    for i in range(6):
        for ax in range(3):
            e = []
            for s in (+1, -1):
                x = x0.copy()
                x[i, ax] += s * h
                o.set_positions(x)
                e.append(o.properties()["Ubond"])
            assert abs(-(e[0] - e[1]) / (2 * h) - f[i, ax]) < 1e-6
    # Newton's third law
    assert np.abs(f.sum(axis=0)).max() < 1e-12


# ---------------------------------------------------------------- integrators
def _free(tmp, integ, diff=1.0, n=4000, box=(30.0, 30.0, 30.0)):
    d = pathlib.Path(tmp)
    rng = np.random.default_rng(5)
    pos = rng.random((n, 3)) * np.asarray(box)
    tild_gen.write_data(d / "input.data", pos, np.ones(n, int), np.arange(1, n + 1), [], box, 1, 3)
    text = tild_gen.deck("input.data", 3, (10, 10, 10), [], integrator=integ, delt=0.01,
                         extra=[f"diffusivity 1 {diff}"])
    return text.replace("bond 1 harmonic 1.5 0.0\n", ""), d


@pytest.mark.parametrize("integ,diff", [("EM", 1.0), ("GJF", 1.0), ("GJF", 2.0)])
def test_free_diffusion(integ, diff, tmp_path):
    text, d = _free(tmp_path, integ, diff)
    o = O.System(text, d, seed=9)
    x0 = o.positions()
    disp = np.zeros_like(x0)
    steps = 200
    L = np.array([30.0, 30.0, 30.0])
    prev = x0
    for _ in range(steps):
        o.run(1)
        x = o.positions()
        disp += U.min_image(x - prev, L)
        prev = x
    msd = (disp ** 2).mean()
    t = steps * 0.01
    # EM: <dx^2> = 2 t (overdamped, unit kT and friction).
    # GJF (device_GJF_integrator.cu:25-48) is underdamped Langevin with friction 1/D and
    # mass m = 1 started at rest: <dx^2> = 2 De t - De tau (3 - 4 e^{-t/tau} + e^{-2t/tau}),
    # tau = m D.  Its noise (sqrt(2 dt)/D) z carries the 1/D factor the reference flags,
    # which makes De = 1 for every D (consistent noise would give De = D).
    if integ == "EM":
        expect = 2 * t
    else:
        tau, De = diff, 1.0
        expect = 2 * De * t - De * tau * (3 - 4 * math.exp(-t / tau) + math.exp(-2 * t / tau))
    assert abs(msd / expect - 1) < 0.05, (msd, expect)


def test_em_wraps_once_into_box(tmp_path):
    text, d = _free(tmp_path, "EM", n=10, box=(5.0, 5.0, 5.0))
    o = O.System(text, d, seed=1)
    x = o.positions()
    x[0] = [4.9999, 0.00001, 2.5]
    o.set_positions(x)
    for _ in range(20):
        o.run(1)
        y = o.positions()
        assert (y >= 0).all() and (y <= 5.0).all()


# ---------------------------------------------------------------- neighbour list
def test_nlist_matches_brute_force(tmp_path):
    text, d = decks.melt3d(tmp_path, n_chains=30, box=(6.0, 7.0, 6.5), grid=(12, 14, 13))
    o = O.System(text, d, seed=1)
    rc = 1.5
    cell_of, counts, neigh, cmax, nc = o.nlist(rc)
    x = o.positions().astype(np.float32)
    L = np.array([o.getd("L0"), o.getd("L1"), o.getd("L2")], dtype=np.float32)
    for i in range(o.ns):
        dxv = x[i] - x
        dxv = np.where(dxv < -L / 2, dxv + L, np.where(dxv > L / 2, dxv - L, dxv))
        ref = np.nonzero((dxv.astype(np.float32) ** 2).sum(axis=1) < np.float32(rc * rc))[0]
        got = np.sort(neigh[i, :counts[i]])
        assert np.array_equal(got, ref), i


PI32 = float(np.float32(3.141592654))


def _field_u(phase, dir_, A, L, r):
    """The field w(r) of pair_style_fieldphases.cu:89-185 (written to fieldphase_map.dat)."""
    a = [2 * PI32 / L[j] for j in range(3)]
    if phase == 0:
        return A * math.sin(a[dir_] * r[dir_])
    if phase == 2:
        d1, d2 = 0, 1
        if dir_ == d1:
            d1 = 2
        elif dir_ == d2:
            d2 = 2
        return A * math.cos(a[d1] * r[d1]) * math.cos(a[d2] * r[d2])
    e = (math.sin(a[0] * r[0]) * math.cos(a[1] * r[1]) + math.sin(a[1] * r[1]) * math.cos(a[2] * r[2])
         + math.sin(a[2] * r[2]) * math.cos(a[0] * r[0]))
    return A * (e * e - 1.44)


@pytest.mark.parametrize("phase,dir_,c", [(0, 1, 1.0), (2, 2, -1.0), (2, 0, -1.0), (3, 0, -0.5)])
def test_fieldphase_forces_follow_the_field(phase, dir_, c, tmp_path):
    # The reference's force tables are c * grad w with c = +1 (lamellar), -1 (cylinders),
    # -1/2 (gyroid).  Its grid force is +rho^2 f on type1 and -rho^2 f on type2, so a bead
    # alone on a node (CIC weight 1, rho = 1/gvol) feels c grad w / gvol, a type-2 bead the
    # opposite, and two beads stacked on one node twice as much (the rho^2 quirk).
    A, L = 0.7, (8.0, 8.0, 8.0)
    nodes = np.array([[1.5, 2.5, 3.0], [4.0, 1.0, 5.5], [6.5, 3.5, 2.0], [6.5, 3.5, 2.0]])
    types = np.array([1, 2, 1, 1])
    tild_gen.write_data(tmp_path / "input.data", nodes, types, np.arange(1, 5), [], L, 2, 3)
    text = tild_gen.deck("input.data", 3, (16, 16, 16), [("fieldphase", 1, 2, A, phase, dir_, 1)])
    o = O.System(text.replace("bond 1 harmonic 1.5 0.0\n", ""), tmp_path, seed=1)
    f = o.forces()
    gvol = 0.5 ** 3
    h = 1e-5
    for i, (mult, sgn) in enumerate([(1, 1), (1, -1), (2, 1), (2, 1)]):
        g = np.zeros(3)
        for ax in range(3):
            rp, rm = nodes[i].copy(), nodes[i].copy()
            rp[ax] += h
            rm[ax] -= h
            g[ax] = (_field_u(phase, dir_, A, L, rp) - _field_u(phase, dir_, A, L, rm)) / (2 * h)
        want = sgn * mult * c * g / gvol
        assert np.allclose(f[i], want, rtol=1e-6, atol=1e-6 * np.abs(want).max()), (i, f[i], want)


def test_fieldphase_bcc_is_rejected(tmp_path):
    text, d = decks.melt3d(tmp_path)
    text += "n_fieldphases 1\nfieldphase 1 2 0.5 1 0 1\n"
    with pytest.raises(ValueError, match="BCC"):
        O.System(text, d, seed=1)


RAMP_STYLES = {
    # (style tuple, its deck line, u per unit prefactor at k = 0)
    "gaussian": (("gaussian", 1, 2, 2.0, 0.5), "gaussian 1 2 2.0 0.5", lambda Rp: 1.0),
    "erf": (("erf", 1, 2, 0.4, 0.6, 0.5), "erf 1 2 0.4 0.6 0.5", lambda Rp: (4.0 * math.pi * Rp ** 3 / 3.0) ** 2),
    "gaussian_erf": (("gaussian_erf", 1, 2, 0.7, 0.6, 0.5), "gaussian_erf 1 2 0.7 0.6 0.5",
                     lambda Rp: 4.0 * math.pi * Rp ** 3 / 3.0),
}


@pytest.mark.parametrize("kind", list(RAMP_STYLES))
def test_ramp_follows_the_reference_update(kind, tmp_path):
    # PairStyle::Update (pair_style.cu:308-328) after step s multiplies the style's
    # tables by new/cur with new = 1000 A0 + delta (s + 1), delta = 1000 (A0 - Af) /
    # max_steps, and cur = 1000 u[k=0], the table's k = 0 value rather than its
    # prefactor.  The amplitude after step s is therefore (A0 + (A0 - Af)(s + 1) /
    # max_steps) / u1(0), with u1(0) the k = 0 value per unit prefactor (1 for a
    # Gaussian, the squared sphere volume for erf, the sphere volume for gaussian_erf):
    # the ramp moves away from Af, and erf styles jump by 1/u1(0) at the first update.
    # The virial tables are never rescaled.  With the beads frozen (delt = 0) the style
    # energy of the log after S steps (table of the update after step S - 1) follows
    # that amplitude exactly, and the pressure stays at its step-0 value.
    style, line, u10 = RAMP_STYLES[kind]
    A0, Af, ms = float(np.float32(style[3])), 5.0, 10    # prefactors are parsed as FP32
    text, d = decks.melt3d(tmp_path, styles=[style], delt=0.0)
    text = text.replace(line, line + f" ramp {Af}").replace("max_steps 1000", f"max_steps {ms}")
    assert f"ramp {Af}" in text and f"max_steps {ms}" in text
    o = O.System(text, d, seed=1)
    x0 = o.positions().copy()
    p0 = o.properties()
    U0 = p0["U_styles"][0]
    Rp = float(np.float32(style[4])) if kind != "gaussian" else 0.0
    for S in range(1, 7):
        o.run(1)
        p = o.properties()
        amp = A0 if S == 1 else (A0 + (A0 - Af) * S / ms) / u10(Rp)
        assert abs(p["U_styles"][0] - U0 * amp / A0) <= 1e-9 * abs(U0 * amp / A0), (kind, S)
        np.testing.assert_allclose(p["Ptens"], p0["Ptens"], rtol=1e-12, atol=1e-15)
    np.testing.assert_array_equal(o.positions(), x0)
