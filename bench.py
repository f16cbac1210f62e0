"""TILD step-engine benchmark on one MI355X.

Headline workload: the large lamellar diblock melt -- 12,800 chains of N = 50 (640,000
beads), box 40x40x80, grid 80x80x160, Gaussian + erf pair styles, GJF dt = 0.01, FP32
particles and fields, FP64 property reductions.  One replica under the engine's launch
policy (runs of 60 steps replayed from a captured hipGraph for small replicas, eager
launches for larger ones such as this); the spatial reorders every 60 steps are folded
into ms_per_step whether or not the timed window holds one.  Synthetic seeded random-walk
chains (no datasets are available offline).

Prints one JSON line: whole-job bead-steps/s, the roofline of the dominant kernel
(algorithmic bytes per launch / mean launch time from HIP events bound to the launches
on the engine's stream), the FP64 oracle timed on the host's own cores (one thread and
all the threads of the host's CPU share), and the FP32-vs-FP64 error of every quantity.
Legs that ride along, one per benchmark configuration:
  C1  24,000-bead diblock melt, 40^3, EM
  C2  the headline melt, also as the "plain" path (eager launches, never a graph)
  C3  81,000-bead coacervate, 64^3, charges + Gaussian-erf, EM, and a sweep of 8
      coacervate replicas batched vs plain
  C4  256 uneven diblock replicas advanced together vs one at a time
  C5  a Poisson stream of jobs through the continuous-batching server, p50/p99
"""
import argparse
import concurrent.futures
import json
import multiprocessing
import os
import pathlib
import platform
import sys
import tempfile
import threading
import time

REPO = pathlib.Path(__file__).resolve().parent
sys.path.insert(0, str(REPO / "gpu-tild_amd"))

HBM_PEAK_GBS = 8000.0
REORDER_EVERY = 60        # automatic spatial-reorder interval of the melts (engine default)


def phase_bytes(ns, D, M, Mh, nf, T, gjf, charged, ncoef, reorder_every, plane_inv=False):
    """HBM bytes of each profiled kernel for one step of one replica.  The particle
    kernel uses the per-unit figures of the reference path's minimal traffic: per bead
    it reads x and the type and writes x (8D + 4 B), plus the GJF state read and written
    (16D B) and the charge (4 B); per grid node it reads each type's gather mask and D
    field components (4T(1 + D) B) and the D electric-field components (4D B).  On top
    of that it writes each bead's 16-byte bucket record (and charge) for the next step.
    The other kernels count this design's minimal bytes (diagnostic only): a bucket
    claim pass (x and type in, record out) runs once per spatial reorder; deposit
    records in, densities out; forward plane densities in, half spectra out; fused pass
    spectra and coupling table in, 2 nf spectra out; y inverse 2 nf spectra in, nf D
    out; x C2R nf D spectra in, 12-byte records out (its records need no density mask
    without field phases).  Where the y inverse and the x C2R run as one plane pass
    (plane_inv), that pass reads the 2 nf spectra of the fused pass and writes the
    records, and is listed under x_c2r."""
    q = 4 if charged else 0
    b = {
        "claim": ns * (4 * D + 4 + 16 + q) / reorder_every,
        "deposit": ns * (16 + q) + nf * M * 4,
        "plane_r2c": nf * M * 4 + nf * Mh * 8,
        "last_fused": nf * Mh * 8 + ncoef * Mh * 4 + 2 * nf * Mh * 8,
        "y_ifft": 2 * nf * Mh * 8 + nf * D * Mh * 8 if D == 3 else 0,
        "x_c2r": nf * D * Mh * 8 + nf * M * 12,
        "particles": ns * (8 * D + 4 + (16 * D if gjf else 0) + q + 16 + q)
                     + M * (4 * T * (1 + D) + (4 * D if charged else 0)),
        "advance": 0,
        "reorder": 0,   # several kernels; not a roofline candidate
    }
    if plane_inv:
        b["y_ifft"] = 0
        b["x_c2r"] = 2 * nf * Mh * 8 + nf * M * 12
    return b


# the kernel of each profiled phase on the default (tiled, forward-plane) path, as named
# in the rocprofv3 summaries
PHASE_KERNEL = {"claim": "k_tile_claim", "deposit": "k_tile_deposit<512>", "plane_r2c": "k_plane_reg<512>",
                "last_fused": "k_last_reg<256, 0>", "y_ifft": "k_grad_reg<128>", "x_c2r": "k_x_c2r<true, 256>",
                "particles": "k_particles<true, true>",
                "advance": "k_advance", "reorder": None}
# with the inverse plane pass (square x-y planes)
PHASE_KERNEL_PLANE_INV = dict(PHASE_KERNEL, y_ifft=None, x_c2r="k_plane_inv<512>")


def step_bytes_minimal(ns, D, M, T, gjf, charged):
    """Whole-step minimal HBM bytes B = ns*b_p + M*b_g: per bead x and type read by the
    spread and the gather and x written (2(4D + 4) + 4D), GJF state (16D), charges read
    twice (8); per node and type the density written, transformed (read + write), read in
    k-space and as the gather mask (20 B) and per (type, direction) the field written and
    gathered (8 B); the charge field adds 16 + 8D B per node."""
    bp = 2 * (4 * D + 4) + 4 * D + (16 * D if gjf else 0) + (8 if charged else 0)
    bg = 4 * (5 * T + 2 * D * T) + (4 * (4 + 2 * D) if charged else 0)
    return ns * bp + M * bg


# ------------------------------------------------------------------ decks
def _gen_sweep(args):
    root, idx, seed = args
    import tild_gen
    d = pathlib.Path(root) / f"r{idx}"
    return tild_gen.sweep_replica(d, idx, seed=seed), str(d)


def _gen_serve(args):
    root, idx, seed = args
    import tild_gen
    d = pathlib.Path(root) / f"j{idx}"
    text, steps = tild_gen.serve_job(d, idx, seed=seed)
    return text, str(d), steps


def _gen_named(args):
    root, name, seed = args
    import tild_gen
    d = pathlib.Path(root) / name
    fn = {"c1": tild_gen.diblock_melt_24k, "c2": tild_gen.lamellar_melt_640k, "c3": tild_gen.coacervate_81k}
    return fn[name.split("_")[0]](d, seed=seed), str(d)


def generate(args, root):
    """Write every deck before the GPU is touched (a pool of fresh writer processes)."""
    ctx = multiprocessing.get_context("spawn")
    with concurrent.futures.ProcessPoolExecutor(max_workers=8, mp_context=ctx) as ex:
        named = ["c1", "c2", "c3"] + [f"c3_{k}" for k in range(args.c3_replicas)]
        fut_named = {n: ex.submit(_gen_named, (root, n, args.seed + (int(n.split("_")[1]) if "_" in n else 0)))
                     for n in named}
        sweep = list(ex.map(_gen_sweep, [(root, i, args.seed) for i in range(args.sweep_replicas)]))
        serve = list(ex.map(_gen_serve, [(root, i, args.seed) for i in range(args.serve_jobs)]))
        decks = {n: f.result() for n, f in fut_named.items()}
    return decks, sweep, serve


# ------------------------------------------------------------------ helpers
def cpu_info():
    model = platform.processor() or platform.machine()
    try:
        for line in pathlib.Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    try:
        share = len(os.sched_getaffinity(0))
    except AttributeError:
        share = os.cpu_count() or 1
    omp = os.environ.get("OMP_NUM_THREADS")
    threads = min(share, int(omp)) if omp and omp.isdigit() and int(omp) > 0 else share
    return model, threads


def _box(s):
    return [s.info("L0"), s.info("L1"), s.info("L2")][:s.dim]


def _rel(a, b):
    import numpy as np
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def fp64_errors(engine_like, oracle, sysm, with_rho=True):
    """FP32 engine against the FP64 oracle at the same positions: densities (peak
    relative), forces (|df| / |f|), energies and pressure (relative)."""
    import numpy as np
    e, o = engine_like, oracle
    f, fo = e.forces(), o.forces()
    out = {}
    if with_rho:
        worst = 0.0
        for t in range(e.nfields):
            ro = o.density(t) if t < sysm.ntypes else o.charge_density()
            worst = max(worst, float(np.abs(e.density(t).astype(np.float64) - ro).max() / np.abs(ro).max()))
        out["rho"] = worst
    out["forces"] = _rel(f, fo)
    p, po = e.properties(), o.properties()
    for k in ("Upe", "Ubond", "Uelec"):
        if po[k] != 0:
            out[k] = abs(p[k] - po[k]) / abs(po[k])
    out["U_styles"] = float(np.max(np.abs(p["U_styles"] - po["U_styles"]) / np.abs(po["U_styles"])))
    out["Ptens"] = float(np.abs(p["Ptens"] - po["Ptens"]).max() / np.abs(po["Ptens"]).max())
    return out


def traj_error(engine, oracle, sysm, n):
    import numpy as np
    engine.step(n)
    oracle.run(n)
    d = engine.positions().astype(np.float64) - oracle.positions()
    L = np.asarray(_box(sysm))
    d -= L * np.round(d / L)
    return float(np.abs(d).max())


def time_oracle(o, steps, threads):
    """Oracle bead-steps/s over `steps` further steps of an oracle system with `threads`
    worker threads (its results do not depend on the count)."""
    sys.path.insert(0, str(REPO / "tests"))
    import oracle_ctypes as O
    O.set_threads(threads)
    t0 = time.perf_counter()
    o.run(steps)
    dt = time.perf_counter() - t0
    return o.ns * steps / dt, dt


def leg_engine(text, d, seed, steps, warmup, graph=True):
    """One replica on its own: `steps` steps after `warmup` under the engine's launch
    policy (graph=True: captured runs of 60 steps for small replicas) or eager launches
    with no graph (the plain path), inputs resident before timing."""
    import tild
    s = tild.System(text, d)
    b = tild.Batch(device=0)
    r = b.add(s, seed)
    b.step(warmup, graph=graph)
    b.properties(r)          # synchronises
    t0 = time.perf_counter()
    b.step(steps, graph=graph)
    b.properties(r)
    wall = time.perf_counter() - t0
    b.close()
    return s.ns * steps / wall, wall / steps * 1e3


# ------------------------------------------------------------------ legs
def leg_headline(args, text, d):
    import numpy as np
    import tild
    sysm = tild.System(text, d)
    ns, D, M = sysm.ns, sysm.dim, sysm.M
    Nx = sysm.Nx
    Mh = (Nx[0] // 2 + 1) * Nx[1] * Nx[2]
    T = sysm.ntypes
    eng = tild.Engine(sysm, device=0, seed=args.seed, reorder_every=args.reorder_every)
    eng.step(args.warmup)
    eng.synchronize()
    start_step = eng.step_count()
    log_every = 1000
    t0 = time.perf_counter()
    done = 0
    nprops = 0
    while done < args.steps:
        n = min(log_every, args.steps - done)
        eng.step(n)
        done += n
        if done % log_every == 0:
            p = eng.properties()
            nprops += 1
            if not np.isfinite(p["Upe"]):
                raise RuntimeError("non-finite energy")
    eng.synchronize()
    wall = time.perf_counter() - t0
    # one property evaluation after the timed window, always
    pend = eng.properties()
    if not np.isfinite(pend["Upe"]):
        raise RuntimeError("non-finite energy after the timed steps")
    # kernel profile over two reorder intervals: phase times per launch, and the
    # reorders (plus the bucket-claim pass that follows each) amortised per step
    prof = eng.profile(args.profile_steps)
    every = args.reorder_every if args.reorder_every > 0 else REORDER_EVERY
    reorder_cost_ms = (prof["reorder"] + prof["claim"]) * args.profile_steps / max(1, args.profile_steps // every)
    in_window = (start_step + args.steps) // every - start_step // every
    correction_ms = (args.steps / every - in_window) * reorder_cost_ms if args.reorder_every >= 0 else 0.0
    ms_per_step = (wall * 1e3 + correction_ms) / args.steps
    value = ns / (ms_per_step * 1e-3)
    plane_inv = bool(eng.geometry()["plane_inv"])
    kernels = PHASE_KERNEL_PLANE_INV if plane_inv else PHASE_KERNEL
    pb = phase_bytes(ns, D, M, Mh, T, T, True, False, T * (T + 1) // 2, every, plane_inv)
    dom = max((k for k in prof if pb[k] > 0), key=lambda k: prof[k])   # dominant kernel
    ach = pb[dom] / (prof[dom] * 1e-3) / 1e9
    step_bytes = step_bytes_minimal(ns, D, M, T, True, False)
    kern_ms = sum(v for k, v in prof.items())
    phases = {k: {"ms": round(v, 5), "GBps": round(pb[k] / (v * 1e-3) / 1e9, 1) if v > 0 and pb[k] else None}
              for k, v in prof.items()}
    eng.close()
    return {
        "value": value, "ms_per_step": ms_per_step, "wall_s": wall, "ns": ns, "Nx": Nx[:D], "props": nprops + 1,
        "Upe_end": pend["Upe"], "reorders_in_window": in_window, "reorder_ms": reorder_cost_ms,
        "timed_ms_per_step": wall * 1e3 / args.steps,
        "roofline": {"bound": "hbm", "kernel": kernels[dom], "achieved": ach, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": ach / HBM_PEAK_GBS, "algorithmic_bytes": pb[dom],
                     "launch_ms": prof[dom]},
        "step_roofline": {"bytes": step_bytes, "kernel_ms": kern_ms,
                          "GBps": step_bytes / (ms_per_step * 1e-3) / 1e9,
                          "frac": step_bytes / (ms_per_step * 1e-3) / 1e9 / HBM_PEAK_GBS},
        "phases": phases, "dominant": dom, "phase_kernels": {k: kernels[k] for k in prof},
    }


def leg_config(args, name, text, d, gjf, charged, steps, err_steps, threads, cpu_steps, cpu_steps_1t):
    """A single-replica configuration: bead-steps/s under the engine's launch policy and
    with plain eager launches, the step's HBM roofline, FP64 errors per quantity, and the
    oracle on 1 and all threads."""
    import tild
    sys.path.insert(0, str(REPO / "tests"))
    import oracle_ctypes as O
    s = tild.System(text, d)
    v_graph, ms_graph = leg_engine(text, d, args.seed, steps, 10, graph=True)
    v_plain, ms_plain = leg_engine(text, d, args.seed, steps, 10, graph=False)
    B = step_bytes_minimal(s.ns, s.dim, s.M, s.ntypes, gjf, charged)
    out = {"beads": s.ns, "grid": s.Nx[:s.dim], "steps": steps,
           "engine": {"value": v_graph, "ms_per_step": ms_graph},
           "plain": {"value": v_plain, "ms_per_step": ms_plain},
           "step_bytes": B, "roofline_frac": B / (ms_graph * 1e-3) / 1e12 / 8.0, "unit": "bead-steps/s"}
    if not args.no_cpu_baseline:
        O.set_threads(threads)
        e = tild.Engine(s, device=0, seed=args.seed)
        o = O.System(text, d, seed=args.seed)
        errs = fp64_errors(e, o, s)
        errs[f"positions_{err_steps}_steps"] = traj_error(e, o, s, err_steps)
        e.close()
        out["fp64_error"] = errs
        vn, tn = time_oracle(o, cpu_steps, threads)
        v1, t1 = time_oracle(o, cpu_steps_1t, 1)
        O.set_threads(threads)
        out["oracle"] = {"1_thread": {"value": v1, "steps": cpu_steps_1t, "s": t1},
                         "all_threads": {"value": vn, "threads": threads, "steps": cpu_steps, "s": tn}}
    return out


def leg_c3_sweep(args, decks):
    """Coacervate replicas (81,000 beads, 64^3 each) advanced together versus one at a
    time with eager launches."""
    import tild
    systems = [tild.System(t, d) for t, d in decks]
    b = tild.Batch(device=0)
    ids = [b.add(s, args.seed + k) for k, s in enumerate(systems)]
    beads = sum(s.ns for s in systems)
    b.step(10, graph=True)
    b.properties(ids[0])
    t0 = time.perf_counter()
    b.step(args.c3_sweep_steps, graph=True)
    b.properties(ids[0])
    tb = time.perf_counter() - t0
    t0 = time.perf_counter()
    for r in ids:
        b.step(args.c3_sweep_steps, replicas=[r], graph=False)
    b.properties(ids[0])
    tp = time.perf_counter() - t0
    b.close()
    bat, pl = beads * args.c3_sweep_steps / tb, beads * args.c3_sweep_steps / tp
    return {"replicas": len(ids), "beads_total": beads, "steps": args.c3_sweep_steps, "value": bat,
            "plain_value": pl, "speedup_vs_plain": bat / pl, "unit": "bead-steps/s"}


def leg_sweep(args, decks):
    """256 uneven replicas: all advanced together (one launch per phase, block geometry
    from the autotuner) versus each advanced alone with eager launches; the step's HBM
    roofline over all replicas; FP64 error of sampled replicas at their final positions."""
    import numpy as np
    import tild
    systems = [tild.System(t, d) for t, d in decks]
    b = tild.Batch(device=0, autotune=not args.no_autotune)
    t0 = time.perf_counter()
    ids = [b.add(s, args.seed + k) for k, s in enumerate(systems)]
    t_add = time.perf_counter() - t0
    geo = [b.geometry(r) for r in ids]
    tuning = {"autotune": not args.no_autotune, "add_s": t_add,
              "classes_measured": sum(1 for g in geo if g["tuned"] and not g["from_cache"]),
              "from_cache": sum(1 for g in geo if g["from_cache"]),
              "sum_tuned_step_ms": sum(g["step_ms"] for g in geo)}
    beads = sum(s.ns for s in systems)
    b.step(5, graph=True)
    b.step(5, graph=False)
    b.properties(ids[0])
    t0 = time.perf_counter()
    b.step(args.sweep_steps, graph=True)
    b.properties(ids[0])
    tb = time.perf_counter() - t0
    t0 = time.perf_counter()
    for r in ids:
        b.step(args.sweep_plain_steps, replicas=[r], graph=False)
    b.properties(ids[0])
    tp = time.perf_counter() - t0
    final = {}
    grids = [s.Nx[0] for s in systems]
    order = np.argsort(grids, kind="stable")
    sample = sorted({int(order[j]) for j in np.linspace(0, len(ids) - 1, args.sweep_oracle_sample).round().astype(int)})
    for k in sample:
        final[k] = b.positions(ids[k])
    b.close()
    errs = []
    if not args.no_cpu_baseline:
        sys.path.insert(0, str(REPO / "tests"))
        import oracle_ctypes as O
        for k in sample:
            t, d = decks[k]
            e = tild.Engine(systems[k], device=0, seed=args.seed + k)
            o = O.System(t, d, seed=args.seed + k)
            e.set_positions(final[k])
            o.set_positions(final[k].astype(np.float64))
            er = fp64_errors(e, o, systems[k], with_rho=False)
            er.update({"replica": k, "grid": grids[k], "beads": systems[k].ns})
            errs.append(er)
            e.close()
    batched = beads * args.sweep_steps / tb
    plain = beads * args.sweep_plain_steps / tp
    # minimal HBM bytes of one step of every replica (EM, uncharged)
    step_bytes = sum(step_bytes_minimal(s.ns, s.dim, s.M, s.ntypes, False, False) for s in systems)
    ms_step = tb / args.sweep_steps * 1e3
    return {"workload": f"{len(ids)} uneven diblock-melt replicas (grids {min(grids)}^3..{max(grids)}^3 at dx=0.5, "
                        f"N in 10..40, A_AB in [1,4], rho0=3, EM)",
            "replicas": len(ids), "beads_total": beads, "steps": args.sweep_steps,
            "value": batched, "unit": "bead-steps/s", "ms_per_step": ms_step,
            "step_bytes": step_bytes, "roofline_frac": step_bytes / (ms_step * 1e-3) / 1e9 / HBM_PEAK_GBS,
            "plain_value": plain, "plain_steps": args.sweep_plain_steps, "speedup_vs_plain": batched / plain,
            "tuning": tuning, "fp64_error_final_state": errs}


def leg_serve(args, jobs, step_work=0, check_oracle=True):
    """A Poisson stream of jobs through the continuous-batching server; FP64 error of
    sampled jobs at their final positions.  step_work > 0: each chunk advances the jobs
    with the least remaining work, up to step_work thousand beads + grid nodes per step."""
    import numpy as np
    import tild
    rng = np.random.default_rng(args.seed)
    gaps = rng.exponential(1.0 / args.serve_rate, size=len(jobs))
    srv = tild.Server(device=0, chunk_steps=args.serve_chunk, use_graph=True, step_work=step_work)
    errors = []
    ids = {}     # job index -> server job id

    def feeder():
        t0 = time.perf_counter()
        at = 0.0
        for k, ((text, d, steps), g) in enumerate(zip(jobs, gaps)):
            at += g
            delay = at - (time.perf_counter() - t0)
            if delay > 0:
                time.sleep(delay)
            try:
                ids[k] = srv.submit(text, d, args.seed, steps)
            except Exception as e:   # reported in the JSON line
                errors.append(str(e))

    def progress():
        st = srv.stats()
        print(f"bench: serving {st['done']}/{len(jobs)} done, {st['running']} running, {st['queued']} queued",
              file=sys.stderr, flush=True)

    t0 = time.perf_counter()
    th = threading.Thread(target=feeder)
    th.start()
    while th.is_alive():
        th.join(timeout=20)
        if th.is_alive():
            progress()
    # the makespan ends when the server has drained (a wait on its own condition, so the
    # clock stops at the last retirement, not at the next poll)
    while True:
        try:
            srv.drain(timeout=20)
            break
        except tild.TildError as e:
            if "timed out" not in str(e) or time.perf_counter() - t0 > 900:
                raise
            progress()
    wall = time.perf_counter() - t0
    st = srv.stats()
    arrivals = float(np.sum(gaps))
    # sampled jobs: FP64 error at their final positions
    errs = []
    if check_oracle and not args.no_cpu_baseline and ids:
        sys.path.insert(0, str(REPO / "tests"))
        import oracle_ctypes as O
        for k in range(0, len(jobs), max(1, len(jobs) // args.serve_oracle_sample)):
            if k not in ids:
                continue
            text, d, steps = jobs[k]
            s = tild.System(text, d)
            if s.ns > args.serve_oracle_max_beads:
                continue
            srv.meta[ids[k]] = (s.ns, s.dim)
            _, x, lat = srv.result(ids[k])
            e = tild.Engine(s, device=0, seed=args.seed)
            o = O.System(text, d, seed=args.seed)
            e.set_positions(x)
            o.set_positions(x.astype(np.float64))
            er = fp64_errors(e, o, s, with_rho=False)
            er.update({"job": k, "dim": s.dim, "grid": s.Nx[:s.dim], "beads": s.ns, "steps": steps,
                       "latency_s": lat})
            errs.append(er)
            e.close()
    srv.close()
    # the worker builds admitted jobs' replicas while a chunk runs, so admission time
    # overlaps device time; what is left is retirement and the gaps between chunks
    idle = wall - st["busy_s"] - st["retire_s"]
    return {"workload": f"{len(jobs)} jobs, Poisson arrivals at {args.serve_rate}/s: 2D films 64^2..256^2 and "
                        f"3D melts 32^3..96^3 (N=20, rho0=3), 1000-5000 steps each",
            "schedule": "every running job each chunk" if step_work <= 0 else
                        f"least remaining work first, {step_work} thousand beads + nodes per step",
            "jobs": len(jobs), "done": st["done"], "failed": st["failed"] + len(errors),
            "p50_s": st["p50_s"], "p99_s": st["p99_s"],
            "queue_wait_p50_s": st["wait50_s"], "queue_wait_p99_s": st["wait99_s"],
            "service_p50_s": st["svc50_s"], "service_p99_s": st["svc99_s"],
            "arrival_span_s": arrivals, "makespan_s": wall,
            "throughput": st["bead_steps"] / wall, "unit": "bead-steps/s",
            "gpu_busy_frac": st["busy_s"] / wall, "admit_overlapped_frac": st["admit_s"] / wall,
            "retire_frac": st["retire_s"] / wall, "idle_frac": idle / wall,
            "peak_hbm_bytes": st["peak_bytes"], "fp64_error_final_state": errs}


def traffic_probe(args):
    """Kernels for the HBM-traffic counter runs: a device-to-device copy of known size
    (the calibration of FETCH_SIZE / WRITE_SIZE) followed by a few eager steps of the
    640,000-bead melt.  Prints one JSON line with the sizes."""
    import ctypes
    import tild
    import tild_gen

    hip = ctypes.CDLL("libamdhip64.so")
    nbytes = 256 << 20
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    for ptr in (a, b):
        if hip.hipMalloc(ctypes.byref(ptr), ctypes.c_size_t(nbytes)) != 0:
            raise RuntimeError("hipMalloc failed")
    for _ in range(3):
        if hip.hipMemcpy(b, a, ctypes.c_size_t(nbytes), 3) != 0:   # hipMemcpyDeviceToDevice
            raise RuntimeError("hipMemcpy failed")
    hip.hipDeviceSynchronize()
    with tempfile.TemporaryDirectory() as td:
        text = tild_gen.lamellar_melt_640k(td, seed=args.seed)
        s = tild.System(text, td)
        e = tild.Engine(s, device=0, seed=args.seed, reorder_every=args.reorder_every)
        for _ in range(args.probe_steps):
            e.step(1)   # one step per call: eager launches, no graph
        e.synchronize()
        out = {"copy_bytes": nbytes, "copies": 3, "steps": args.probe_steps, "beads": s.ns, "M": s.M}
        e.close()
    hip.hipFree(a)
    hip.hipFree(b)
    print(json.dumps(out))


def log(msg):
    print(f"bench: {msg}", file=sys.stderr, flush=True)


def run(args):
    import tild  # noqa: F401  (fails loudly if the library is missing)

    model, threads = cpu_info()
    with tempfile.TemporaryDirectory() as td:
        t0 = time.perf_counter()
        decks, sweep_decks, serve_jobs = generate(args, td)
        log(f"decks written in {time.perf_counter() - t0:.1f} s")
        text, mdir = decks["c2"]
        head = leg_headline(args, text, mdir)
        log(f"C2 headline {head['value']:.3e} bead-steps/s")
        v_plain, ms_plain = leg_engine(text, mdir, args.seed, args.plain_steps, 10, graph=False)
        log(f"C2 plain {v_plain:.3e}")

        cpu, c2_err = None, None
        if not args.no_cpu_baseline:
            sys.path.insert(0, str(REPO / "tests"))
            import oracle_ctypes as O
            import tild as T
            O.set_threads(threads)
            s2 = T.System(text, mdir)
            e2 = T.Engine(s2, device=0, seed=args.seed)
            o2 = O.System(text, mdir, seed=args.seed)
            c2_err = fp64_errors(e2, o2, s2)
            c2_err["positions_3_steps"] = traj_error(e2, o2, s2, 3)
            e2.close()
            log("C2 FP64 errors done")
            vn, tn = time_oracle(o2, args.cpu_steps, threads)
            v1, t1 = time_oracle(o2, args.cpu_steps_1t, 1)
            O.set_threads(threads)
            del o2
            cpu = {"value": vn, "unit": "bead-steps/s", "cores": threads, "kind": "port",
                   "sample": f"{args.cpu_steps} steps of the same 640,000-bead deck (FP64 oracle, OpenMP on "
                             f"{threads} threads of {model}), {tn:.1f} s",
                   "single_thread": {"value": v1, "steps": args.cpu_steps_1t, "s": t1, "cores": 1},
                   "cpu_model": model}
            log("C2 CPU baseline done")

        configs = {}
        if not args.skip_configs:
            t1, d1 = decks["c1"]
            configs["C1"] = leg_config(args, "C1", t1, d1, False, False, args.c1_steps, 20, threads, 20, 20)
            log("C1 done")
            t3, d3 = decks["c3"]
            configs["C3"] = leg_config(args, "C3", t3, d3, False, True, args.c3_steps, 20, threads, 20, 3)
            log("C3 done")
            if args.c3_replicas > 0:
                configs["C3"]["sweep"] = leg_c3_sweep(args, [decks[f"c3_{k}"] for k in range(args.c3_replicas)])
                log("C3 sweep done")
        batched = leg_sweep(args, sweep_decks) if sweep_decks else None
        log("C4 done")
        serving = leg_serve(args, serve_jobs) if serve_jobs else None
        log("C5 done")
        if serve_jobs and args.serve_step_work > 0:
            serving["least_work_first"] = leg_serve(args, serve_jobs, step_work=args.serve_step_work,
                                                    check_oracle=False)
            log("C5 (least remaining work first) done")

    traffic = None
    tfile = REPO / "profiles" / "traffic_melt640k.json"
    if tfile.exists():
        try:
            traffic = json.loads(tfile.read_text())["per_kernel_bytes"].get(head["roofline"]["kernel"])
        except (ValueError, OSError, KeyError):
            traffic = None
    roof = dict(head["roofline"])
    roof["traffic"] = traffic
    configs["C2"] = {"beads": head["ns"], "grid": head["Nx"],
                     "engine": {"value": head["value"], "ms_per_step": head["ms_per_step"]},
                     "plain": {"value": v_plain, "ms_per_step": ms_plain},
                     "roofline_frac": head["step_roofline"]["frac"], "fp64_error": c2_err}
    out = {
        "metric": "bead-steps/s per GPU + HBM roofline %; batched vs plain; p50/p99 sweep latency",
        "value": head["value"],
        "unit": "bead-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": head["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "lamellar diblock melt, 12800 x N=50 = 640000 beads, box 40x40x80, "
                               "grid 80x80x160, 3 gaussian + 3 erf styles, GJF dt=0.01, single replica, "
                               "engine launch policy (eager at this size); spatial reorders every 60 steps "
                               "amortised in",
                   "beads": head["ns"], "grid": head["Nx"], "property_evaluations": head["props"],
                   "timed_ms_per_step": head["timed_ms_per_step"], "reorders_in_window": head["reorders_in_window"],
                   "reorder_ms": head["reorder_ms"], "Upe_after": head["Upe_end"]},
        "roofline": roof,
        "step_roofline": head["step_roofline"],
        "phases": head["phases"],
        "plain": {"value": v_plain, "ms_per_step": ms_plain, "steps": args.plain_steps,
                  "note": "one replica, eager launches, no graph"},
        "oracle_error": c2_err,
        "configs": configs,
        "batched": batched,
        "serving": serving,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--profile-steps", type=int, default=120, help="two reorder intervals")
    ap.add_argument("--plain-steps", type=int, default=200)
    ap.add_argument("--cpu-steps", type=int, default=20, help="oracle steps on all threads (headline deck)")
    ap.add_argument("--cpu-steps-1t", type=int, default=3, help="oracle steps on one thread (headline deck)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--skip-configs", action="store_true", help="skip the C1 and C3 legs")
    ap.add_argument("--c1-steps", type=int, default=500)
    ap.add_argument("--c3-steps", type=int, default=500)
    ap.add_argument("--c3-replicas", type=int, default=8, help="coacervate sweep size (0 skips it)")
    ap.add_argument("--c3-sweep-steps", type=int, default=100)
    ap.add_argument("--traffic-probe", action="store_true",
                    help="only run the kernels of the HBM-traffic counter runs (copy calibration + eager steps)")
    ap.add_argument("--probe-steps", type=int, default=5)
    ap.add_argument("--sweep-replicas", type=int, default=256, help="0 skips the batched-sweep leg")
    ap.add_argument("--sweep-steps", type=int, default=200)
    ap.add_argument("--sweep-plain-steps", type=int, default=20)
    ap.add_argument("--sweep-oracle-sample", type=int, default=4)
    ap.add_argument("--no-autotune", action="store_true", help="batched sweep without the block-geometry autotuner")
    ap.add_argument("--serve-jobs", type=int, default=500, help="0 skips the served-stream leg")
    ap.add_argument("--serve-rate", type=float, default=20.0)
    ap.add_argument("--serve-chunk", type=int, default=100)
    ap.add_argument("--serve-oracle-sample", type=int, default=8)
    ap.add_argument("--serve-step-work", type=int, default=6000,
                    help="second pass of the stream with least-remaining-work-first chunks of this many "
                         "thousand beads + grid nodes per step (0 skips it)")
    ap.add_argument("--serve-oracle-max-beads", type=int, default=120000)
    ap.add_argument("--reorder-every", type=int, default=0,
                    help="spatial reorder interval (0 automatic, n > 0 every n steps, < 0 never)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py: this engine runs on one GPU only (independent replicas shard across GPUs "
              "without communication); --gpus must be 1", file=sys.stderr)
        sys.exit(2)
    if args.traffic_probe:
        traffic_probe(args)
    else:
        run(args)


if __name__ == "__main__":
    main()
