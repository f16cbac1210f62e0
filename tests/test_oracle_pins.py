"""CPU tier: closed-form and equilibrium pins of the FP64 oracle's integrators and of
its 2D (and on-node 3D) energy, force and pressure paths.

The integrator drift terms (device_EM_integrator.cu:24-26, device_GJF_integrator.cu:44-47)
are checked with the thermal noise switched off against explicit Euler and against the
closed-form solution of the GJF recurrence for a harmonic dimer, and with the noise on
against the stationary bond-length variance.  The 2D paths are checked against the
continuum pair potential, which a grid reproduces exactly for beads on grid nodes."""
import math
import pathlib

import numpy as np
import pytest

import decks
import oracle_ctypes as O
import tild_gen


def _dimers(tmp, n_dimers, integ, k, r0, delt, box=20.0, mass=1.0, diff=None, seed=3, sep=1.0):
    """n_dimers bonded pairs (u = k (r - r0)^2, bonds.cu:28) and no pair styles."""
    d = pathlib.Path(tmp)
    d.mkdir(parents=True, exist_ok=True)
    rng = np.random.default_rng(seed)
    a = box * (0.2 + 0.6 * rng.random((n_dimers, 3)))
    u = rng.normal(size=(n_dimers, 3))
    u /= np.linalg.norm(u, axis=1, keepdims=True)
    b = a + sep * (0.6 + 0.8 * rng.random((n_dimers, 1))) * u
    pos = np.stack([a, b], axis=1).reshape(-1, 3)
    bonds = [(2 * i + 1, 2 * i + 2) for i in range(n_dimers)]
    tild_gen.write_data(d / "input.data", pos, np.ones(2 * n_dimers, int), np.repeat(np.arange(1, n_dimers + 1), 2),
                        bonds, (box, box, box), 1, 3, masses=[mass])
    extra = [f"diffusivity 1 {diff}"] if diff is not None else []
    text = tild_gen.deck("input.data", 3, (8, 8, 8), [], integrator=integ, delt=delt, bond_k=k, bond_r0=r0,
                         extra=extra, max_steps=100000)
    return text, d


def f32(v):
    """Deck scalars are parsed into FP32 (read_input.cu, config_utils.cu)."""
    return float(np.float32(v))


def _bond_forces(x, k, r0, L):
    """-grad of k (r - r0)^2 for consecutive pairs (0,1), (2,3), ... (minimum image)."""
    dr = x[0::2] - x[1::2]
    dr -= L * np.round(dr / L)
    r = np.linalg.norm(dr, axis=1, keepdims=True)
    f1 = -2.0 * k * (r - r0) / r * dr
    f = np.empty_like(x)
    f[0::2], f[1::2] = f1, -f1
    return f


def _unwrap_pairs(x, L):
    dr = x[0::2] - x[1::2]
    return dr - L * np.round(dr / L)


# ---------------------------------------------------------------- zero noise
def test_em_without_noise_is_explicit_euler(tmp_path):
    # one EM step with the noise off is x + delt f(x) (device_EM_integrator.cu:24-26)
    k, r0, dt, L = 1.7, 0.4, 0.01, 20.0
    text, d = _dimers(tmp_path, 6, "EM", k, r0, dt)
    o = O.System(text, d, seed=1)
    o.set_noise(0.0)
    x0 = o.positions()
    o.run(1)
    want = x0 + f32(dt) * _bond_forces(x0, f32(k), f32(r0), L)
    np.testing.assert_allclose(o.positions(), want, rtol=0, atol=1e-12)


def test_em_harmonic_dimer_decays_geometrically(tmp_path):
    # u = k r^2: the bond vector obeys d' = (1 - 4 k delt) d, the centre of mass is still
    k, dt, L, n = 1.5, 0.01, 20.0, 40
    text, d = _dimers(tmp_path, 5, "EM", k, 0.0, dt)
    o = O.System(text, d, seed=1)
    o.set_noise(0.0)
    x0 = o.positions()
    d0 = _unwrap_pairs(x0, L)
    com0 = 0.5 * (x0[0::2] + x0[1::2])
    o.run(n)
    x = o.positions()
    np.testing.assert_allclose(_unwrap_pairs(x, L), d0 * (1 - 4 * k * f32(dt)) ** n, rtol=1e-10, atol=1e-12)
    np.testing.assert_allclose(0.5 * (x[0::2] + x[1::2]), com0, atol=1e-12)


@pytest.mark.parametrize("mass,diff", [(1.0, 1.0), (2.0, 0.5), (0.7, 3.0)])
def test_gjf_harmonic_dimer_matches_closed_form(mass, diff, tmp_path):
    # GJF in its Verlet form (Gronbech-Jensen & Farago, Mol. Phys. 111, 983 (2013)) with
    # friction alpha = kT / D = 1 / D (device_GJF_integrator.cu:25-48):
    #   x_{n+1} = 2 b x_n - a x_{n-1} + b dt^2 / m f_n,  b = 1 / (1 + alpha dt / 2m),
    #   a = (1 - alpha dt / 2m) b, and x_{-1} = x_0 (device_comm_utils.cu:15-16).
    # For u = k r^2 the bond vector sees f = -4 k d, so d_n = c1 l1^n + c2 l2^n with
    # l1, l2 the roots of l^2 - (2b - 4 k b dt^2 / m) l + a = 0 and d_{-1} = d_0.
    k, dt, L, n = 1.5, 0.02, 20.0, 60
    text, d = _dimers(tmp_path, 4, "GJF", k, 0.0, dt, mass=mass, diff=diff)
    o = O.System(text, d, seed=1)
    o.set_noise(0.0)
    x0 = o.positions()
    d0 = _unwrap_pairs(x0, L)
    com0 = 0.5 * (x0[0::2] + x0[1::2])
    dt, mass, alpha = f32(dt), f32(mass), 1.0 / f32(diff)
    b = 1.0 / (1.0 + alpha * dt / (2 * mass))
    a = (1.0 - alpha * dt / (2 * mass)) * b
    l1, l2 = np.roots([1.0, -(2 * b - 4 * k * b * dt * dt / mass), a]).astype(complex)
    # c1 + c2 = 1 (n = 0), c1 / l1 + c2 / l2 = 1 (n = -1)
    c1, c2 = np.linalg.solve(np.array([[1, 1], [1 / l1, 1 / l2]]), np.array([1.0, 1.0]))
    # the bond force is cut to zero below r^2 = 1e-4 (device_bonds.cu:41-47); compare
    # over the steps before any bond first gets shorter than 0.05
    amps = np.array([(c1 * l1 ** j + c2 * l2 ** j).real for j in range(n + 1)])
    short = np.nonzero(np.abs(amps) * np.linalg.norm(d0, axis=1).min() < 0.05)[0]
    n = int(short[0]) - 1 if len(short) else n
    assert n >= 20
    traj = []
    for step in range(1, n + 1):
        o.run(1)
        if step % 5 == 0 or step == n:
            traj.append((step, o.positions()))
    for step, x in traj:
        amp = (c1 * l1 ** step + c2 * l2 ** step).real
        np.testing.assert_allclose(_unwrap_pairs(x, L), d0 * amp, rtol=1e-9, atol=1e-11)
        np.testing.assert_allclose(0.5 * (x[0::2] + x[1::2]), com0, atol=1e-11)
    # the bond has moved well away from its start over the window
    assert abs(amps[n] - 1.0) > 0.1


# ---------------------------------------------------------------- equilibrium
def _mean_r2(o, L, burn, nsamp, every):
    o.run(burn)
    acc = []
    for _ in range(nsamp):
        o.run(every)
        acc.append((_unwrap_pairs(o.positions(), L) ** 2).sum(axis=1).mean())
    return float(np.mean(acc))


def test_em_dimer_bond_variance_is_the_discrete_ou_value(tmp_path):
    # u = k r^2 with kT = 1: per component the bond vector follows the discrete
    # Ornstein-Uhlenbeck recurrence d' = (1 - 4 k dt) d + sqrt(2 dt)(z1 - z2), whose
    # stationary variance is 4 dt / (1 - (1 - 4 k dt)^2) -- above the Boltzmann value
    # 1 / (2k) by O(dt), which the test resolves.
    k, dt, L = 1.5, 0.02, 20.0
    text, d = _dimers(tmp_path, 1500, "EM", k, 0.0, dt)
    o = O.System(text, d, seed=7)
    got = _mean_r2(o, L, burn=100, nsamp=60, every=15)
    dt = f32(dt)
    want = 3 * 4 * dt / (1 - (1 - 4 * k * dt) ** 2)
    boltz = 3 / (2 * k)
    assert abs(got / want - 1) < 0.015, (got, want)
    assert abs(got / boltz - 1) > 0.03, (got, boltz)


def test_gjf_dimer_bond_variance_is_boltzmann(tmp_path):
    # GJF samples a harmonic potential's configurational distribution exactly at any
    # stable dt (Gronbech-Jensen & Farago 2013): <r^2> = 3 / (2k) at D = 1, where the
    # reference's noise amplitude sqrt(2 dt) / D (device_GJF_integrator.cu:35-36) is the
    # consistent sqrt(2 alpha kT dt).
    k, dt, L = 1.5, 0.05, 20.0
    text, d = _dimers(tmp_path, 1500, "GJF", k, 0.0, dt, diff=1.0)
    o = O.System(text, d, seed=7)
    got = _mean_r2(o, L, burn=100, nsamp=60, every=15)
    want = 3 / (2 * k)
    assert abs(got / want - 1) < 0.015, (got, want)


# ---------------------------------------------------------------- beads on grid nodes
def _on_node_pair(tmp, dim, style, box, grid, r1, r2):
    d = pathlib.Path(tmp)
    d.mkdir(parents=True, exist_ok=True)
    pos = np.array([r1, r2], dtype=np.float64)
    bx = tuple(box) + ((1.0,) if dim == 2 else ())
    tild_gen.write_data(d / "input.data", pos, np.array([1, 2]), np.array([1, 2]), [], bx, 2, dim)
    text = tild_gen.deck("input.data", dim, grid, [style])
    return text.replace("bond 1 harmonic 1.5 0.0\n", ""), d


@pytest.mark.parametrize("dim", [2, 3])
def test_gaussian_pair_on_nodes_is_the_continuum_pair(dim, tmp_path):
    # For u(k) = A exp(-k^2 s^2 / 2) (pair_style_gaussian.cu:72-90) the pair potential is
    # u(r) = A (2 pi s^2)^{-D/2} exp(-r^2 / 2 s^2).  Two beads of different type on grid
    # nodes (CIC weight exactly 1; dx a power of two) see the grid's image sum of u, which
    # for s << L is u itself: the style energy (pair_style.cu:204-233) is u(r12), the
    # forces (pair_style.cu:81-196 + gather) are -/+ grad u, and the virial pressure
    # (pair_style.cu:235-276) is r12_a F_b / (V D) with F the force on bead 1 (the
    # off-diagonal component is dr_b f_a, device_array_utils.cu:477-491).
    A, s = 1.75, 0.875    # exact in FP32, the parser's precision
    box = (16.0, 16.0, 16.0)[:dim]
    grid = (64, 64, 64)[:dim]
    r1 = np.array([5.25, 7.5, 6.0])[:dim]
    r2 = np.array([6.5, 6.75, 7.25])[:dim]
    text, d = _on_node_pair(tmp_path, dim, ("gaussian", 1, 2, A, s), box, grid, r1, r2)
    o = O.System(text, d, seed=1)
    f = o.forces()
    p = o.properties()
    r = r1 - r2
    rr = float(np.linalg.norm(r))
    u = A * (2 * math.pi * s * s) ** (-dim / 2) * math.exp(-rr * rr / (2 * s * s))
    F = u * r / (s * s)                      # -grad_1 u
    V = float(np.prod(box))
    assert abs(p["U_styles"][0] / u - 1) < 1e-10
    np.testing.assert_allclose(f[0], F, rtol=1e-9, atol=1e-12 * np.abs(F).max())
    np.testing.assert_allclose(f[1], -F, rtol=1e-9, atol=1e-12 * np.abs(F).max())
    if dim == 2:
        want = np.array([r[0] * F[0], r[1] * F[1], r[1] * F[0]]) / (V * dim)
        got = p["Ptens"][:3]
    else:
        want = np.array([r[0] * F[0], r[1] * F[1], r[2] * F[2], r[1] * F[0], r[2] * F[0], r[2] * F[1]]) / (V * dim)
        got = p["Ptens"]
    np.testing.assert_allclose(got, want, rtol=1e-9, atol=1e-12 * np.abs(want).max())


def test_two_particle_gaussian_force_2d_converges_off_node(tmp_path):
    # Off the nodes CIC + ik differentiation approach -grad u at second order in dx.
    A, s = 2.0, 1.0
    errs = []
    for n in (40, 80):
        text, d = _on_node_pair(tmp_path / str(n), 2, ("gaussian", 1, 2, A, s), (20.0, 20.0), (n, n),
                                [10.05, 10.13], [11.27, 10.41])
        o = O.System(text, d, seed=1)
        f = o.forces()
        x = o.positions()
        r = x[0] - x[1]
        rr = np.linalg.norm(r)
        u = A / (2 * math.pi * s * s) * math.exp(-rr ** 2 / (2 * s * s))
        fa = u * r / (s * s)
        errs.append(np.linalg.norm(f[0] - fa) / np.linalg.norm(fa))
        assert np.linalg.norm(f[0] + f[1]) < 1e-9 * np.linalg.norm(fa)
    assert errs[1] < 0.05 and errs[0] / errs[1] > 3.0, errs


@pytest.mark.parametrize("P", [0, 1, 2, 3])
def test_spread_sum_rule_2d(P, tmp_path):
    text, d = decks.melt2d(tmp_path, pmeorder=P)
    o = O.System(text, d, seed=1)
    o.forces()
    gv = o.getd("gvol")
    nt = np.bincount(o.types(), minlength=o.ntypes)
    for t in range(o.ntypes):
        assert abs(o.density(t).sum() * gv - nt[t]) < 1e-9 * nt[t]
    _, W = o.stencil()
    np.testing.assert_allclose(W.sum(axis=1) * gv, 1.0, atol=1e-12)


def _cardinal_bspline(k, t):
    if k == 1:
        return 1.0 if 0.0 <= t < 1.0 else 0.0
    return (t * _cardinal_bspline(k - 1, t) + (k - t) * _cardinal_bspline(k - 1, t - 1.0)) / (k - 1)


@pytest.mark.parametrize("P", [0, 1, 2, 3])
def test_spread_weights_2d_are_cardinal_bsplines(P, tmp_path):
    # 2D stencils are the (P+1)^2 products of the per-axis cardinal B-splines M_{P+1},
    # ordered x-outer, y-inner (device_grid_utils.cu:153-226)
    text, d = decks.melt2d(tmp_path, n_chains=8, pmeorder=P)
    o = O.System(text, d, seed=1)
    o.forces()
    _, W = o.stencil()
    gv = o.getd("gvol")
    x = o.positions().astype(np.float32)
    dxs = [np.float32(o.getd(f"dx{a}")) for a in range(2)]
    worst = 0.0
    for i in range(o.ns):
        w1 = []
        for a in range(2):
            xf, dxf = x[i, a], dxs[a]
            if P % 2 == 0:
                g = int(np.float32(xf + np.float32(0.5) * dxf) / dxf)
                sx = (float(xf) - g * float(dxf)) / float(dxf)
            else:
                g = int(np.float32(xf / dxf))
                sx = (float(xf) - (g + 0.5) * float(dxf)) / float(dxf)
            w1.append([_cardinal_bspline(P + 1, 0.5 + j - sx) for j in range(P + 1)])
        ref = np.outer(w1[0], w1[1]).ravel()
        worst = max(worst, np.abs(W[i] * gv - ref).max())
    assert worst < 1e-12
