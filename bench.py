"""TILD step-engine benchmark on one MI355X.

Workload (default): the large lamellar diblock melt -- 12,800 chains of N = 50
(640,000 beads), box 40x40x80, grid 80x80x160, Gaussian + erf pair styles, GJF
dt = 0.01, FP32 particles and fields, FP64 property reductions every 1,000 steps.
Synthetic seeded random-walk chains (no datasets are available offline).

Prints one JSON line: whole-job bead-steps/s, the roofline of the dominant kernel
(algorithmic bytes per launch / mean launch time from HIP events on the engine's
stream) and the FP64 CPU oracle timed on a bounded sample of the same workload.
Two more legs ride along: a sweep of 256 uneven replicas advanced together versus
one at a time ("batched"), and a Poisson stream of jobs through the job server with
p50/p99 time to result ("serving").
"""
import argparse
import concurrent.futures
import json
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


def phase_bytes(ns, D, M, Mh, nf, T, gjf, charged, ncoef, reorder_every):
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
    out; x C2R nf D spectra and the densities in, 12-byte records out."""
    q = 4 if charged else 0
    b = {
        "claim": ns * (4 * D + 4 + 16 + q) / reorder_every,
        "deposit": ns * (16 + q) + nf * M * 4,
        "plane_r2c": nf * M * 4 + nf * Mh * 8,
        "last_fused": nf * Mh * 8 + ncoef * Mh * 4 + 2 * nf * Mh * 8,
        "y_ifft": 2 * nf * Mh * 8 + nf * D * Mh * 8 if D == 3 else 0,
        "x_c2r": nf * D * Mh * 8 + T * M * 4 + nf * M * 12,
        "particles": ns * (8 * D + 4 + (16 * D if gjf else 0) + q + 16 + q)
                     + M * (4 * T * (1 + D) + (4 * D if charged else 0)),
        "advance": 0,
        "reorder": 0,   # several kernels; not a roofline candidate
    }
    return b


# the kernel of each profiled phase on the default (tiled, forward-plane) path, as named
# in the rocprofv3 summaries
PHASE_KERNEL = {"claim": "k_tile_claim", "deposit": "k_tile_deposit<512>", "plane_r2c": "k_plane_r2c<512>",
                "last_fused": "k_last_fused<256>", "y_ifft": "k_mid_grad<256>", "x_c2r": "k_x_c2r<true, 256>",
                "particles": "k_particles<true, true>",
                "advance": "k_advance", "reorder": None}


def step_bytes_minimal(ns, D, M, T, gjf, charged):
    """Whole-step minimal HBM bytes B = ns*b_p + M*b_g: per bead x and type read by the
    spread and the gather and x written (2(4D + 4) + 4D), GJF state (16D), charges read
    twice (8); per node and type the density written, transformed (read + write), read in
    k-space and as the gather mask (20 B) and per (type, direction) the field written and
    gathered (8 B); the charge field adds 16 + 8D B per node."""
    bp = 2 * (4 * D + 4) + 4 * D + (16 * D if gjf else 0) + (8 if charged else 0)
    bg = 4 * (5 * T + 2 * D * T) + (4 * (4 + 2 * D) if charged else 0)
    return ns * bp + M * bg


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


def generate(args, root):
    """Write every deck before the GPU is touched (a pool of forked writers)."""
    import tild_gen
    main_dir = pathlib.Path(root) / "main"
    main_text = tild_gen.lamellar_melt_640k(main_dir, seed=args.seed)
    sweep, serve = [], []
    with concurrent.futures.ProcessPoolExecutor(max_workers=8) as ex:
        if args.sweep_replicas > 0:
            sweep = list(ex.map(_gen_sweep, [(root, i, args.seed) for i in range(args.sweep_replicas)]))
        if args.serve_jobs > 0:
            serve = list(ex.map(_gen_serve, [(root, i, args.seed) for i in range(args.serve_jobs)]))
    return (main_text, str(main_dir)), sweep, serve


def leg_sweep(args, decks):
    """256 uneven replicas: all advanced together (one launch per phase, graph replay)
    versus each advanced alone with eager launches."""
    import tild
    systems = [tild.System(t, d) for t, d in decks]
    b = tild.Batch(device=0)
    ids = [b.add(s, args.seed + k) for k, s in enumerate(systems)]
    beads = sum(s.ns for s in systems)
    b.step(5, graph=True)
    b.step(5, graph=False)
    t0 = time.perf_counter()
    b.step(args.sweep_steps, graph=True)
    tb = time.perf_counter() - t0
    t0 = time.perf_counter()
    for r in ids:
        b.step(args.sweep_plain_steps, replicas=[r], graph=False)
    tp = time.perf_counter() - t0
    b.close()
    batched = beads * args.sweep_steps / tb
    plain = beads * args.sweep_plain_steps / tp
    grids = sorted({s.Nx[0] for s in systems})
    return {"workload": f"{len(ids)} uneven diblock-melt replicas (grids {grids[0]}^3..{grids[-1]}^3 at dx=0.5, "
                        f"N in 10..40, A_AB in [1,4], rho0=3, EM)",
            "replicas": len(ids), "beads_total": beads, "steps": args.sweep_steps,
            "value": batched, "unit": "bead-steps/s", "ms_per_step": tb / args.sweep_steps * 1e3,
            "plain_value": plain, "plain_steps": args.sweep_plain_steps, "speedup_vs_plain": batched / plain}


def leg_serve(args, jobs):
    """A Poisson stream of jobs through the continuous-batching server."""
    import numpy as np
    import tild
    rng = np.random.default_rng(args.seed)
    gaps = rng.exponential(1.0 / args.serve_rate, size=len(jobs))
    srv = tild.Server(device=0, chunk_steps=args.serve_chunk, use_graph=True)
    errors = []

    def feeder():
        t0 = time.perf_counter()
        at = 0.0
        for (text, d, steps), g in zip(jobs, gaps):
            at += g
            delay = at - (time.perf_counter() - t0)
            if delay > 0:
                time.sleep(delay)
            try:
                srv.submit(text, d, args.seed, steps)
            except Exception as e:   # reported in the JSON line
                errors.append(str(e))

    t0 = time.perf_counter()
    th = threading.Thread(target=feeder)
    th.start()
    th.join()
    srv.drain(timeout=600)
    wall = time.perf_counter() - t0
    st = srv.stats()
    srv.close()
    return {"workload": f"{len(jobs)} jobs, Poisson arrivals at {args.serve_rate}/s: 2D films 64^2..256^2 and "
                        f"3D melts 32^3..96^3 (N=20, rho0=3), 1000-5000 steps each",
            "jobs": len(jobs), "done": st["done"], "failed": st["failed"] + len(errors),
            "p50_s": st["p50_s"], "p99_s": st["p99_s"], "makespan_s": wall,
            "throughput": st["bead_steps"] / wall, "unit": "bead-steps/s",
            "gpu_busy_frac": st["busy_s"] / wall, "peak_hbm_bytes": st["peak_bytes"]}


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


def run(args):
    import numpy as np
    import tild
    import tild_gen

    with tempfile.TemporaryDirectory() as td:
        (text, mdir), sweep_decks, serve_jobs = generate(args, td)
        sysm = tild.System(text, mdir)
        ns, D, M = sysm.ns, sysm.dim, sysm.M
        Nx = sysm.Nx
        Mh = (Nx[0] // 2 + 1) * Nx[1] * Nx[2]
        T = sysm.ntypes
        eng = tild.Engine(sysm, device=0, seed=args.seed, reorder_every=args.reorder_every)

        eng.step(args.warmup)
        eng.synchronize()
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
        t1 = time.perf_counter()
        wall = t1 - t0
        value = ns * args.steps / wall

        prof = eng.profile(args.profile_steps)
        # automatic reorder interval of this workload: 60 steps
        pb = phase_bytes(ns, D, M, Mh, T, T, True, False, T * (T + 1) // 2,
                         args.reorder_every if args.reorder_every > 0 else (60 if args.reorder_every == 0 else args.steps))
        dom = max((k for k in prof if pb[k] > 0), key=lambda k: prof[k])   # dominant kernel
        ach = pb[dom] / (prof[dom] * 1e-3) / 1e9
        step_bytes = step_bytes_minimal(ns, D, M, T, True, False)
        step_ms = sum(prof.values())
        phases = {k: {"ms": round(v, 5), "GBps": round(pb[k] / (v * 1e-3) / 1e9, 1) if v > 0 and pb[k] else None}
                  for k, v in prof.items()}
        eng.close()

        batched = leg_sweep(args, sweep_decks) if sweep_decks else None
        serving = leg_serve(args, serve_jobs) if serve_jobs else None

        cpu = None
        if not args.no_cpu_baseline:
            sys.path.insert(0, str(REPO / "tests"))
            import oracle_ctypes as O
            osys = O.System(text, mdir, seed=args.seed)
            osys.run(1)
            tc0 = time.perf_counter()
            osys.run(args.cpu_steps)
            tc = time.perf_counter() - tc0
            cpu = {"value": ns * args.cpu_steps / tc, "unit": "bead-steps/s", "cores": 1, "kind": "port",
                   "sample": f"{args.cpu_steps} steps of the same 640,000-bead deck (FP64 oracle, 1 thread, "
                             f"{platform.processor() or platform.machine()}), {tc:.1f} s"}

    traffic = None
    tfile = REPO / "profiles" / "traffic_melt640k.json"
    if tfile.exists():
        try:
            traffic = json.loads(tfile.read_text())["per_kernel_bytes"].get(PHASE_KERNEL[dom])
        except (ValueError, OSError, KeyError):
            traffic = None
    out = {
        "metric": "bead-steps/s per GPU + HBM roofline %; batched vs plain; p50/p99 sweep latency",
        "value": value,
        "unit": "bead-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "lamellar diblock melt, 12800 x N=50 = 640000 beads, box 40x40x80, "
                               "grid 80x80x160, 3 gaussian + 3 erf styles, GJF dt=0.01, plain single replica",
                   "beads": ns, "grid": Nx[:D], "property_evaluations": nprops},
        "roofline": {"bound": "hbm", "kernel": PHASE_KERNEL[dom], "achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": ach / HBM_PEAK_GBS, "traffic": traffic,
                     "algorithmic_bytes": pb[dom], "launch_ms": prof[dom]},
        "step_roofline": {"bytes": step_bytes, "ms": step_ms, "GBps": step_bytes / (step_ms * 1e-3) / 1e9,
                          "frac": step_bytes / (step_ms * 1e-3) / 1e9 / HBM_PEAK_GBS},
        "phases": phases,
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
    ap.add_argument("--profile-steps", type=int, default=20)
    ap.add_argument("--cpu-steps", type=int, default=3)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--traffic-probe", action="store_true",
                    help="only run the kernels of the HBM-traffic counter runs (copy calibration + eager steps)")
    ap.add_argument("--probe-steps", type=int, default=5)
    ap.add_argument("--sweep-replicas", type=int, default=256, help="0 skips the batched-sweep leg")
    ap.add_argument("--sweep-steps", type=int, default=200)
    ap.add_argument("--sweep-plain-steps", type=int, default=20)
    ap.add_argument("--serve-jobs", type=int, default=48, help="0 skips the served-stream leg")
    ap.add_argument("--serve-rate", type=float, default=20.0)
    ap.add_argument("--serve-chunk", type=int, default=100)
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
