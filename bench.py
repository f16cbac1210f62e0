"""TILD step-engine benchmark on one MI355X.

Workload (default): the large lamellar diblock melt -- 12,800 chains of N = 50
(640,000 beads), box 40x40x80, grid 80x80x160, Gaussian + erf pair styles, GJF
dt = 0.01, FP32 particles and fields, FP64 property reductions every 1,000 steps.
Synthetic seeded random-walk chains (no datasets are available offline).

Prints one JSON line: whole-job bead-steps/s, the roofline of the dominant kernel
(algorithmic bytes per launch / mean launch time from HIP events on the engine's
stream) and the FP64 CPU oracle timed on a bounded sample of the same workload.
"""
import argparse
import json
import os
import pathlib
import platform
import sys
import tempfile
import time

REPO = pathlib.Path(__file__).resolve().parent
sys.path.insert(0, str(REPO / "gpu-tild_amd"))

HBM_PEAK_GBS = 8000.0


def phase_bytes(ns, D, M, Mh, nf, T, gjf, charged, ncoef):
    """Algorithmic (minimal unique) HBM bytes of each phase for one step of one replica."""
    part_in = ns * (4 * D + 4 + (4 if charged else 0))
    b = {
        "spread": part_in + nf * M * 4,
        "x_r2c": nf * M * 4 + nf * Mh * 8,
        "y_fft": 2 * nf * Mh * 8 if D == 3 else 0,
        "last_fused": nf * Mh * 8 + ncoef * Mh * 4 + nf * D * Mh * 8,
        "y_ifft": 2 * nf * D * Mh * 8 if D == 3 else 0,
        "x_c2r": nf * D * Mh * 8 + T * M * 4 + nf * M * 16,
        "particles": part_in + ns * (24 + 1 + 4 * D) + nf * M * 16 + (ns * 8 * D if gjf else 0),
        "advance": 0,
    }
    return b


def run(args):
    import numpy as np
    import tild
    import tild_gen

    with tempfile.TemporaryDirectory() as td:
        text = tild_gen.lamellar_melt_640k(td, seed=args.seed)
        sysm = tild.System(text, td)
        ns, D, M = sysm.ns, sysm.dim, sysm.M
        Nx = sysm.Nx
        Mh = (Nx[0] // 2 + 1) * Nx[1] * Nx[2]
        T = sysm.ntypes
        eng = tild.Engine(sysm, device=0, seed=args.seed)

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
        pb = phase_bytes(ns, D, M, Mh, T, T, True, False, T * (T + 1) // 2)
        dom = max(prof, key=lambda k: prof[k])
        ach = pb[dom] / (prof[dom] * 1e-3) / 1e9
        step_bytes = sum(pb.values())
        step_ms = sum(prof.values())
        phases = {k: {"ms": round(v, 5), "GBps": round(pb[k] / (v * 1e-3) / 1e9, 1) if v > 0 else None}
                  for k, v in prof.items()}
        eng.close()

        cpu = None
        if not args.no_cpu_baseline:
            sys.path.insert(0, str(REPO / "tests"))
            import oracle_ctypes as O
            osys = O.System(text, td, seed=args.seed)
            osys.run(1)
            tc0 = time.perf_counter()
            osys.run(args.cpu_steps)
            tc = time.perf_counter() - tc0
            cpu = {"value": ns * args.cpu_steps / tc, "unit": "bead-steps/s", "cores": 1, "kind": "port",
                   "sample": f"{args.cpu_steps} steps of the same 640,000-bead deck (FP64 oracle, 1 thread, "
                             f"{platform.processor() or platform.machine()}), {tc:.1f} s"}

    traffic = None
    tfile = REPO / "profiles" / "traffic_c2.json"
    if tfile.exists():
        try:
            traffic = json.loads(tfile.read_text()).get(dom)
        except (ValueError, OSError):
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
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": ach / HBM_PEAK_GBS, "traffic": traffic,
                     "algorithmic_bytes": pb[dom], "launch_ms": prof[dom]},
        "step_roofline": {"bytes": step_bytes, "ms": step_ms, "GBps": step_bytes / (step_ms * 1e-3) / 1e9},
        "phases": phases,
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
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py: this engine runs on one GPU only (independent replicas shard across GPUs "
              "without communication); --gpus must be 1", file=sys.stderr)
        sys.exit(2)
    run(args)


if __name__ == "__main__":
    main()
