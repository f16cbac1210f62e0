"""Seeded generator of TILD input decks and data files in the reference formats.

The reference ships no generator (its examples name an external tool), so sweeps,
benchmarks and tests build their systems here: chains are random walks with Gaussian
steps of per-axis variance b^2/3 starting uniformly in the box, written in the
LAMMPS-like data format of config_utils.cu:5-364 (charged decks add a q column).
"""
import math
import pathlib

import numpy as np


def _walks(rng, n_chains, N, box, dim, b=1.0):
    box = np.asarray(box[:dim], dtype=np.float64)
    start = rng.random((n_chains, 1, dim)) * box
    steps = rng.normal(0.0, b / math.sqrt(3.0), size=(n_chains, N - 1, dim))
    pos = np.concatenate([start, start + np.cumsum(steps, axis=1)], axis=1)
    return np.mod(pos, box)


def write_data(path, pos, types, mols, bonds, box, ntypes, dim, charges=None, nbond_types=1, masses=None):
    """Write a data file; pos (ns, dim), types/mols 1-based, bonds list of (i, j) 1-based,
    masses per type (default 1)."""
    ns = len(pos)
    lines = ["TILD synthetic configuration", "", f"{ns} atoms", f"{len(bonds)} bonds", "0 angles", "",
             f"{ntypes} atom types", f"{nbond_types} bond types", "0 angle types", ""]
    zhi = box[2] if dim == 3 else 1.0
    lines += [f"0.000000 {box[0]:.6f} xlo xhi", f"0.000000 {box[1]:.6f} ylo yhi",
              f"0.000000 {zhi:.6f} zlo zhi", "", "Masses", ""]
    mass = [1.0] * ntypes if masses is None else list(masses)
    lines += [f"{t + 1} {mass[t]:.6f}" for t in range(ntypes)]
    lines += ["", "Atoms", ""]
    out = "\n".join(lines) + "\n"
    coords = pos if dim == 3 else np.concatenate([pos, np.zeros((ns, 1))], axis=1)
    rows = []
    for i in range(ns):
        c = coords[i]
        if charges is None:
            rows.append(f"{i + 1} {mols[i]} {types[i]}  {c[0]:.6f} {c[1]:.6f} {c[2]:.6f}")
        else:
            rows.append(f"{i + 1} {mols[i]} {types[i]}  {charges[i]:.3f}  {c[0]:.6f} {c[1]:.6f} {c[2]:.6f}")
    out += "\n".join(rows) + "\n\nBonds\n\n"
    out += "\n".join(f"{k + 1} 1  {a} {b}" for k, (a, b) in enumerate(bonds)) + "\n"
    pathlib.Path(path).write_text(out)


def diblock(rng, n_chains, N, fA, box, dim):
    pos = _walks(rng, n_chains, N, box, dim).reshape(-1, dim)
    nA = int(round(fA * N))
    types = np.tile(np.where(np.arange(N) < nA, 1, 2), n_chains)
    mols = np.repeat(np.arange(1, n_chains + 1), N)
    bonds = [(c * N + k + 1, c * N + k + 2) for c in range(n_chains) for k in range(N - 1)]
    return pos, types, mols, bonds


def deck(data_name, dim, grid, styles, integrator="EM", delt=0.005, seed=1, max_steps=1000,
         log_freq=1000, charges=None, pmeorder=1, bond_k=1.5, bond_r0=0.0, extra=()):
    """Input deck text.  styles: list of (kind, t1, t2, A, params...); a fieldphase
    style is ("fieldphase", t1, t2, A, phase, dir, n_periods)."""
    L = [f"Dim {dim}", f"max_steps {max_steps}", f"log_freq {log_freq}", "binary_freq 0",
         f"pmeorder {pmeorder}", f"delt {delt}", f"rand_seed {seed}"]
    if charges is not None:
        L.append(f"charges {charges[0]} {charges[1]}")
    L += [f"read_data {data_name}", f"integrator all {integrator}",
          f"Nx {grid[0]}", f"Ny {grid[1]}"] + ([f"Nz {grid[2]}"] if dim == 3 else [])
    L.append(f"bond 1 harmonic {bond_k} {bond_r0}")
    L += list(extra)
    kinds = {"gaussian": [], "erf": [], "gaussian_erf": [], "fieldphase": []}
    for s in styles:
        kinds[s[0]].append(s)
    for k, lst in kinds.items():
        if lst:
            L.append(f"n_{k}s {len(lst)}")
            for s in lst:
                L.append(" ".join([k] + [str(v) for v in s[1:]]))
    return "\n".join(L) + "\n"


def melt(dirpath, n_chains, N, box, grid, fA=0.5, dim=3, styles=None, integrator="EM", delt=0.005,
         seed=1, max_steps=1000, pmeorder=1, extra=()):
    """Diblock melt deck + data in dirpath; returns the deck text."""
    dirpath = pathlib.Path(dirpath)
    dirpath.mkdir(parents=True, exist_ok=True)
    rng = np.random.default_rng(seed)
    pos, types, mols, bonds = diblock(rng, n_chains, N, fA, box, dim)
    write_data(dirpath / "input.data", pos, types, mols, bonds, box, 2, dim)
    if styles is None:
        styles = [("gaussian", 1, 1, 1.0, 0.5), ("gaussian", 2, 2, 1.0, 0.5), ("gaussian", 1, 2, 2.0, 0.5)]
    text = deck("input.data", dim, grid, styles, integrator, delt, seed, max_steps, max_steps,
                pmeorder=pmeorder, extra=extra)
    (dirpath / "input").write_text(text)
    return text


def coacervate(dirpath, n_chains_each=1000, N=25, n_counter_pairs=2500, total=81000, box=(30, 30, 30),
               grid=(64, 64, 64), seed=1, delt=0.005, lB=2.0, smear=0.5, Rp=0.5, A=2.0, sigma=0.5,
               max_steps=2000):
    """Polycation/polyanion chains + counterions + solvent, four types, q column."""
    dirpath = pathlib.Path(dirpath)
    dirpath.mkdir(parents=True, exist_ok=True)
    rng = np.random.default_rng(seed)
    dim = 3
    pc = _walks(rng, n_chains_each, N, box, dim).reshape(-1, dim)
    pa = _walks(rng, n_chains_each, N, box, dim).reshape(-1, dim)
    npoly = 2 * n_chains_each * N
    nsolv = total - npoly - 2 * n_counter_pairs
    if nsolv < 0:
        raise ValueError("too many beads for the requested total")
    ions = rng.random((2 * n_counter_pairs, dim)) * np.asarray(box)
    solv = rng.random((nsolv, dim)) * np.asarray(box)
    pos = np.concatenate([pc, pa, ions, solv])
    types = np.concatenate([np.full(len(pc), 1), np.full(len(pa), 2), np.full(2 * n_counter_pairs, 3),
                            np.full(nsolv, 4)])
    q = np.concatenate([np.ones(len(pc)), -np.ones(len(pa)), np.ones(n_counter_pairs),
                        -np.ones(n_counter_pairs), np.zeros(nsolv)])
    nchain = 2 * n_chains_each
    mols = np.concatenate([np.repeat(np.arange(1, nchain + 1), N),
                           np.arange(nchain + 1, nchain + 1 + 2 * n_counter_pairs + nsolv)])
    bonds = [(c * N + k + 1, c * N + k + 2) for c in range(nchain) for k in range(N - 1)]
    write_data(dirpath / "input.data", pos, types, mols, bonds, box, 4, dim, charges=q)
    styles = [("gaussian_erf", i, j, A, Rp, sigma) for i in range(1, 5) for j in range(i, 5)]
    text = deck("input.data", dim, grid, styles, "EM", delt, seed, max_steps, max_steps, charges=(lB, smear))
    (dirpath / "input").write_text(text)
    return text


# ------------------------------------------------------------------ named workloads
def diblock_melt_24k(dirpath, seed=1):
    """1,200 AB diblocks of N=20 (24,000 beads) in a 20^3 box, 40^3 grid, Gaussians, EM."""
    return melt(dirpath, 1200, 20, (20.0, 20.0, 20.0), (40, 40, 40), seed=seed, max_steps=1000)


def lamellar_styles():
    return [("gaussian", 1, 1, 1.0, 0.5), ("gaussian", 2, 2, 1.0, 0.5), ("gaussian", 1, 2, 2.0, 0.5),
            ("erf", 1, 1, 0.05, 1.0, 0.5), ("erf", 2, 2, 0.05, 1.0, 0.5), ("erf", 1, 2, 0.1, 1.0, 0.5)]


def lamellar_melt_640k(dirpath, seed=1):
    """12,800 AB diblocks of N=50 (640,000 beads), box 40x40x80, grid 80x80x160,
    Gaussian + erf styles, GJF dt = 0.01."""
    return melt(dirpath, 12800, 50, (40.0, 40.0, 80.0), (80, 80, 160), styles=lamellar_styles(),
                integrator="GJF", delt=0.01, seed=seed, max_steps=2000)


def coacervate_81k(dirpath, seed=1):
    """Polycation + polyanion (2 x 1,000 chains of N=25), 2,500 counterion pairs and solvent
    to 81,000 beads in a 30^3 box, 64^3 grid, smeared charges + Gaussian-erf, EM."""
    return coacervate(dirpath, seed=seed)


SMOOTH_GRIDS = (24, 25, 27, 30, 32, 36, 40, 45, 48)


def sweep_replica(dirpath, idx, seed=1):
    """One replica of the uneven diblock sweep: grid from the 5-smooth set at dx = 0.5,
    N in {10,20,30,40}, A_AB in [1, 4], rho0 = 3, EM."""
    rng = np.random.default_rng(seed * 100003 + idx)
    n = int(rng.choice(SMOOTH_GRIDS))
    L = 0.5 * n
    N = int(rng.choice([10, 20, 30, 40]))
    aab = float(rng.uniform(1.0, 4.0))
    nch = max(1, int(round(3.0 * L ** 3 / N)))
    styles = [("gaussian", 1, 1, 1.0, 0.5), ("gaussian", 2, 2, 1.0, 0.5), ("gaussian", 1, 2, round(aab, 4), 0.5)]
    return melt(dirpath, nch, N, (L, L, L), (n, n, n), styles=styles, seed=seed * 1000 + idx, max_steps=2000)


SERVE_GRIDS_3D = (32, 40, 48, 64, 80, 96)
SERVE_GRIDS_2D = (64, 96, 128, 192, 256)


def serve_job(dirpath, idx, seed=1):
    """One job of the served stream: half 2D diblock films on 64^2..256^2 grids, half 3D
    melts on 32^3..96^3 grids (dx = 0.5, rho0 = 3, N = 20, 12k-330k beads), with
    1,000-5,000 steps.  Returns (deck text, steps)."""
    rng = np.random.default_rng(seed * 7919 + idx)
    steps = int(rng.integers(1000, 5001))
    styles = [("gaussian", 1, 1, 1.0, 0.5), ("gaussian", 2, 2, 1.0, 0.5),
              ("gaussian", 1, 2, round(float(rng.uniform(1.0, 4.0)), 4), 0.5)]
    if idx % 2 == 0:
        n = int(rng.choice(SERVE_GRIDS_2D))
        L = 0.5 * n
        nch = max(1, int(round(3.0 * L * L / 20)))
        text = melt(dirpath, nch, 20, (L, L, 1.0), (n, n), dim=2, styles=styles, seed=seed * 1000 + idx,
                    max_steps=steps)
    else:
        n = int(rng.choice(SERVE_GRIDS_3D))
        L = 0.5 * n
        nch = max(1, int(round(3.0 * L ** 3 / 20)))
        text = melt(dirpath, nch, 20, (L, L, L), (n, n, n), styles=styles, seed=seed * 1000 + idx,
                    max_steps=steps)
    return text, steps


# ------------------------------------------------------------------ command line
def main(argv=None):
    """Write an initial configuration (`input.data`) and its input deck (`input`) into a
    directory, ready for `tild-run -in input` run from there:

        python tild_gen.py melt OUT --chains 400 --N 20 --box 10 10 10 --grid 20 20 20 \\
            --fA 0.5 --chi 2.0 --integrator GJF --delt 0.01 --max-steps 5000 --seed 7
        python tild_gen.py melt OUT --dim 2 --chains 300 --N 20 --box 20 20 --grid 40 40
        python tild_gen.py coacervate OUT --chains-each 100 --N 25 --counter-pairs 250 \\
            --total 8100 --box 14 14 14 --grid 28 28 28
        python tild_gen.py lamellar640k OUT

    A melt is AB diblocks (random-walk chains, fraction fA of A) with Gaussian styles
    A-A and B-B of prefactor 1 and A-B of prefactor --chi (width 0.5)."""
    import argparse
    ap = argparse.ArgumentParser(prog="tild_gen", description=main.__doc__.split("\n\n")[0])
    sub = ap.add_subparsers(dest="kind", required=True)
    m = sub.add_parser("melt")
    m.add_argument("out")
    m.add_argument("--chains", type=int, required=True)
    m.add_argument("--N", type=int, required=True)
    m.add_argument("--box", type=float, nargs="+", required=True)
    m.add_argument("--grid", type=int, nargs="+", required=True)
    m.add_argument("--dim", type=int, default=3, choices=(2, 3))
    m.add_argument("--fA", type=float, default=0.5)
    m.add_argument("--chi", type=float, default=2.0)
    m.add_argument("--integrator", default="EM", choices=("EM", "GJF"))
    m.add_argument("--delt", type=float, default=0.005)
    m.add_argument("--max-steps", type=int, default=1000)
    m.add_argument("--pmeorder", type=int, default=1)
    m.add_argument("--seed", type=int, default=1)
    c = sub.add_parser("coacervate")
    c.add_argument("out")
    c.add_argument("--chains-each", type=int, default=100)
    c.add_argument("--N", type=int, default=25)
    c.add_argument("--counter-pairs", type=int, default=250)
    c.add_argument("--total", type=int, default=8100)
    c.add_argument("--box", type=float, nargs=3, default=(14.0, 14.0, 14.0))
    c.add_argument("--grid", type=int, nargs=3, default=(28, 28, 28))
    c.add_argument("--max-steps", type=int, default=2000)
    c.add_argument("--seed", type=int, default=1)
    w = sub.add_parser("lamellar640k")
    w.add_argument("out")
    w.add_argument("--seed", type=int, default=1)
    a = ap.parse_args(argv)
    if a.kind == "melt":
        if len(a.box) != a.dim or len(a.grid) != a.dim:
            ap.error("--box and --grid take one value per dimension")
        box = tuple(a.box) + ((1.0,) if a.dim == 2 else ())
        styles = [("gaussian", 1, 1, 1.0, 0.5), ("gaussian", 2, 2, 1.0, 0.5), ("gaussian", 1, 2, a.chi, 0.5)]
        melt(a.out, a.chains, a.N, box, tuple(a.grid), fA=a.fA, dim=a.dim, styles=styles,
             integrator=a.integrator, delt=a.delt, seed=a.seed, max_steps=a.max_steps, pmeorder=a.pmeorder)
    elif a.kind == "coacervate":
        coacervate(a.out, a.chains_each, a.N, a.counter_pairs, a.total, tuple(a.box), tuple(a.grid),
                   seed=a.seed, max_steps=a.max_steps)
    else:
        lamellar_melt_640k(a.out, seed=a.seed)
    print(f"wrote {pathlib.Path(a.out) / 'input'} and {pathlib.Path(a.out) / 'input.data'}")
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
