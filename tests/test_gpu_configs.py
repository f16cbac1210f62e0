"""GPU parity against the FP64 oracle at the sizes of the benchmarked workloads.

  C1  1,200 diblocks of N = 20 (24,000 beads), box 20^3, grid 40^3, Gaussians, EM
  C2  12,800 diblocks of N = 50 (640,000 beads), box 40x40x80, grid 80x80x160,
      three Gaussian + three erf styles, GJF dt = 0.01
  C3  polycation/polyanion chains, counterions and solvent (81,000 beads), box 30^3,
      grid 64^3, smeared charges + Gaussian-erf, EM, cell lists at rc = 1.5
  C4  256 uneven diblock replicas (grids 24^3..48^3) advanced as one batch
  C5  a Poisson stream of 2D films and 3D melts through the job server

Each test prints the error of every quantity against the oracle.  Tolerances (FP32
engine against FP64): densities 2e-6 of the peak, forces 1e-5 (|df| / |f|), energies
and pressure 1e-5 relative, trajectories 1e-4 absolute in position."""
import concurrent.futures
import multiprocessing
import pathlib
import time

import numpy as np
import pytest

import gpu_util as U
import oracle_ctypes as O
import tild
import tild_gen

pytestmark = pytest.mark.gpu


def _box(s):
    return [s.info("L0"), s.info("L1"), s.info("L2")][:s.dim]


def _check_forces_and_properties(tag, s, e, o):
    f = e.forces()
    fo = o.forces()
    errs = {}
    for t in range(e.nfields):
        ro = o.density(t) if t < s.ntypes else o.charge_density()
        errs[f"rho{t}"] = float(np.abs(e.density(t).astype(np.float64) - ro).max() / np.abs(ro).max())
    errs["forces"] = U.rel_norm(f, fo)
    p, po = e.properties(), o.properties()
    for k in ("Upe", "Ubond", "Uelec"):
        errs[k] = abs(p[k] - po[k]) / max(abs(po[k]), 1e-300) if po[k] != 0 else abs(p[k])
    errs["U_styles"] = float(np.max(np.abs(p["U_styles"] - po["U_styles"]) / np.abs(po["U_styles"])))
    errs["Ptens"] = float(np.abs(p["Ptens"] - po["Ptens"]).max() / np.abs(po["Ptens"]).max())
    print(f"[{tag}] ns={s.ns} grid={s.grid_shape} vs FP64: " + ", ".join(f"{k} {v:.2e}" for k, v in errs.items()),
          flush=True)
    for t in range(e.nfields):
        assert errs[f"rho{t}"] < 2e-6, (tag, t)
    assert errs["forces"] < 1e-5, tag
    for k in ("Upe", "Ubond", "Uelec", "U_styles", "Ptens"):
        assert errs[k] < 1e-5, (tag, k)


def _check_trajectory(tag, s, e, o, n):
    e.step(n)
    o.run(n)
    dx = U.min_image(e.positions().astype(np.float64) - o.positions(), _box(s))
    err = float(np.abs(dx).max())
    print(f"[{tag}] {n}-step trajectory vs FP64: max |dx| {err:.2e}", flush=True)
    assert err < 1e-4, tag
    return err


def test_c1_diblock_melt_24k(tmp_path):
    text = tild_gen.diblock_melt_24k(tmp_path, seed=1)
    s = tild.System(text, tmp_path)
    assert s.ns == 24000 and s.grid_shape == (40, 40, 40)
    e = tild.Engine(s, device=0, seed=1)
    o = O.System(text, tmp_path, seed=1)
    _check_forces_and_properties("C1", s, e, o)
    _check_trajectory("C1 EM", s, e, o, 20)
    e.close()


def test_c2_lamellar_melt_640k(tmp_path):
    # the benchmarked workload: tiled spread, forward plane pass and the compile-time
    # 10x8 / 16x10 two-stage transforms of the 80- and 160-point axes
    text = tild_gen.lamellar_melt_640k(tmp_path, seed=1)
    s = tild.System(text, tmp_path)
    assert s.ns == 640000 and s.grid_shape == (160, 80, 80)
    e = tild.Engine(s, device=0, seed=1)
    o = O.System(text, tmp_path, seed=1)
    _check_forces_and_properties("C2", s, e, o)
    _check_trajectory("C2 GJF", s, e, o, 3)
    e.close()


def test_c3_coacervate_81k(tmp_path):
    text = tild_gen.coacervate_81k(tmp_path, seed=1)
    s = tild.System(text, tmp_path)
    assert s.ns == 81000 and s.grid_shape == (64, 64, 64) and s.info("do_charges") == 1
    e = tild.Engine(s, device=0, seed=1)
    o = O.System(text, tmp_path, seed=1)
    _check_forces_and_properties("C3", s, e, o)
    # cell lists at rc = 1.5: 20^3 cells; lists are longer than the reference's
    # CELL_MAX here (it would write past them), so they get more room
    rc = 1.5
    cell_of, counts, neigh, cmax, nc = e.nlist(rc, cap=4096)
    ocell, ocounts, oneigh, ocmax, onc = o.nlist(rc)
    assert (cmax, nc) == (ocmax, onc) == (75, 8000)
    assert np.array_equal(cell_of, ocell) and np.array_equal(counts, ocounts)
    for i in range(s.ns):
        assert np.array_equal(neigh[i, :counts[i]], oneigh[i, :counts[i]]), i
    print(f"[C3] cell lists rc={rc}: {nc} cells, CELL_MAX {cmax}, lists {counts.mean():.1f} mean / "
          f"{counts.max()} max, bit-exact", flush=True)
    with pytest.raises(tild.TildError) as ei:
        e.nlist(rc)                      # room for CELL_MAX only: the reference's overrun
    assert ei.value.status == 2
    _check_trajectory("C3 EM", s, e, o, 20)
    e.close()


def _gen_sweep(args):
    root, idx, seed = args
    d = pathlib.Path(root) / f"r{idx}"
    return tild_gen.sweep_replica(d, idx, seed=seed), str(d)


def _spawn_pool():
    # decks are written by fresh interpreters (no copy of this process's device state)
    return concurrent.futures.ProcessPoolExecutor(max_workers=8, mp_context=multiprocessing.get_context("spawn"))


def test_c4_sweep_256_replicas(tmp_path):
    n, steps, seed = 256, 20, 1
    with _spawn_pool() as ex:
        decks = list(ex.map(_gen_sweep, [(str(tmp_path), i, seed) for i in range(n)]))
    systems = [tild.System(t, d) for t, d in decks]
    grids = [s.Nx[0] for s in systems]
    print(f"[C4] {n} replicas, grids {min(grids)}^3..{max(grids)}^3, {sum(s.ns for s in systems)} beads", flush=True)
    b = tild.Batch(device=0)
    ids = [b.add(s, seed + k) for k, s in enumerate(systems)]
    b.step(10, graph=True)
    b.step(10, graph=True)
    xb = [b.positions(r) for r in ids]
    pb = [b.properties(r) for r in ids]
    b.close()
    # every replica against itself run alone
    for k, s in enumerate(systems):
        e = tild.Engine(s, device=0, seed=seed + k)
        e.step(steps)
        assert np.array_equal(e.positions(), xb[k]), k
        e.close()
    print(f"[C4] all {n} batched replicas bit-identical to runs alone", flush=True)
    # eight replicas spanning the grid sizes against the oracle
    order = np.argsort(grids, kind="stable")
    pick = sorted({int(order[j]) for j in np.linspace(0, n - 1, 8).round().astype(int)})
    for k in pick:
        t, d = decks[k]
        o = O.System(t, d, seed=seed + k)
        o.run(steps)
        s = systems[k]
        err = float(np.abs(U.min_image(xb[k].astype(np.float64) - o.positions(), _box(s))).max())
        po = o.properties()
        eU = abs(pb[k]["Upe"] - po["Upe"]) / abs(po["Upe"])
        eP = float(np.abs(pb[k]["Ptens"] - po["Ptens"]).max() / np.abs(po["Ptens"]).max())
        print(f"[C4] replica {k}: grid {s.Nx[0]}^3, {s.ns} beads, {steps} steps vs FP64: max |dx| {err:.2e}, "
              f"Upe {eU:.2e}, Ptens {eP:.2e}", flush=True)
        assert err < 1e-4 and eU < 1e-5 and eP < 1e-5, k


def _gen_job(args):
    root, idx, seed = args
    d = pathlib.Path(root) / f"j{idx}"
    text, _ = tild_gen.serve_job(d, idx, seed=seed)
    return text, str(d)


def test_c5_served_stream(tmp_path):
    # jobs of the served stream's shapes (2D films 64^2..256^2, 3D melts 32^3..96^3),
    # shortened to 20-60 steps, submitted at Poisson times; every job equals its run
    # alone bit for bit and the smaller ones are checked against the oracle
    nj, seed = 16, 1
    with _spawn_pool() as ex:
        jobs = list(ex.map(_gen_job, [(str(tmp_path), i, seed) for i in range(nj)]))
    systems = [tild.System(t, d) for t, d in jobs]
    rng = np.random.default_rng(5)
    steps = [int(v) for v in rng.integers(20, 61, size=nj)]
    gaps = rng.exponential(1.0 / 20.0, size=nj)
    srv = tild.Server(device=0, chunk_steps=10, use_graph=True)
    ids = []
    for (t, d), s, n, g in zip(jobs, systems, steps, gaps):
        time.sleep(float(g))
        ids.append(srv.submit(t, d, seed, n, ns=s.ns, dim=s.dim))
    srv.drain(timeout=300)
    st = srv.stats()
    assert st["done"] == nj and st["failed"] == 0
    res = [srv.result(j) for j in ids]
    srv.close()
    for k, (s, n) in enumerate(zip(systems, steps)):
        e = tild.Engine(s, device=0, seed=seed)
        e.step(n)
        assert np.array_equal(e.positions(), res[k][1]), k
        assert e.properties()["Upe"] == res[k][0]["Upe"], k
        e.close()
    print(f"[C5] {nj} jobs ({sum(s.ns for s in systems)} beads): results bit-identical alone and in the mix; "
          f"p50 {st['p50_s']:.3f} s p99 {st['p99_s']:.3f} s", flush=True)
    small = sorted(range(nj), key=lambda k: systems[k].ns)
    sample = [k for k in small if systems[k].dim == 2][:2] + [k for k in small if systems[k].dim == 3][:2]
    for k in sample:
        t, d = jobs[k]
        o = O.System(t, d, seed=seed)
        o.run(steps[k])
        s = systems[k]
        err = float(np.abs(U.min_image(res[k][1].astype(np.float64) - o.positions(), _box(s))).max())
        po = o.properties()
        eU = abs(res[k][0]["Upe"] - po["Upe"]) / abs(po["Upe"])
        print(f"[C5] job {k}: {s.dim}D grid {s.grid_shape}, {s.ns} beads, {steps[k]} steps vs FP64: "
              f"max |dx| {err:.2e}, Upe {eU:.2e}", flush=True)
        assert err < 1e-4 and eU < 1e-5, k
