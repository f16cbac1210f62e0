"""Job server: a job's result is the same alone and in the mix; admission honours the budget."""
import numpy as np
import pytest

import decks
import tild

pytestmark = pytest.mark.gpu


def _jobs(tmp):
    specs = [
        decks.melt3d(tmp / "a", n_chains=40, N=10, box=(6.0, 6.0, 6.0), grid=(12, 12, 12)),
        decks.melt3d(tmp / "b", n_chains=30, N=20, box=(7.5, 7.5, 9.0), grid=(15, 15, 18),
                     integrator="GJF", delt=0.01, styles=decks.MIXED),
        decks.melt2d(tmp / "c"),
        decks.charged(tmp / "d"),
        decks.melt3d(tmp / "e", n_chains=20, N=10, box=(6.0, 5.0, 7.0), grid=(12, 10, 14), pmeorder=2),
        decks.melt3d(tmp / "f", n_chains=40, N=10, box=(6.0, 6.0, 6.0), grid=(12, 12, 12), seed=9),
    ]
    steps = [5, 12, 7, 3, 9, 11]
    return [(t, d, 40 + i, n) for i, ((t, d), n) in enumerate(zip(specs, steps))]


def _alone(jobs):
    out = []
    for text, d, seed, n in jobs:
        s = tild.System(text, d)
        e = tild.Engine(s, device=0, seed=seed)
        e.step(n)
        out.append((e.positions(), e.properties(), s.ns, s.dim))
        e.close()
    return out


@pytest.mark.parametrize("graph,step_work", [(True, 0), (False, 0), (True, 3)],
                         ids=["graph", "eager", "least-work-first"])
def test_jobs_same_alone_and_in_the_mix(graph, step_work, tmp_path):
    # step_work 3: each chunk advances only the jobs with the least remaining work, about
    # 3,000 beads + grid nodes per step (one or two of these jobs)
    jobs = _jobs(tmp_path)
    ref = _alone(jobs)
    srv = tild.Server(device=0, chunk_steps=4, max_replicas=3, use_graph=graph, step_work=step_work)
    ids = [srv.submit(t, d, seed, n, ns=r[2], dim=r[3]) for (t, d, seed, n), r in zip(jobs, ref)]
    srv.drain(timeout=120)
    for j, (x0, p0, _, _) in zip(ids, ref):
        assert srv.poll(j) == "done"
        p, x, lat = srv.result(j)
        assert np.array_equal(x, x0), j
        assert p["Upe"] == p0["Upe"] and p["step"] == p0["step"], j
        assert lat > 0
    st = srv.stats()
    assert st["done"] == len(jobs) and st["failed"] == 0 and st["queued"] == 0 and st["running"] == 0
    assert st["bytes_in_use"] == 0
    assert 0 < st["p50_s"] <= st["p99_s"]
    assert st["bead_steps"] == sum(r[2] * n for r, (_, _, _, n) in zip(ref, jobs))
    print(f"server (graph={graph}, step_work={step_work}): {len(jobs)} jobs bit-identical to runs alone; "
          f"p50 {st['p50_s']:.3f} s, "
          f"p99 {st['p99_s']:.3f} s")
    srv.close()


def test_admission_within_budget(tmp_path):
    jobs = _jobs(tmp_path)
    # room for about two of the small jobs at a time
    srv = tild.Server(device=0, chunk_steps=3, use_graph=True, hbm_budget=5 << 20)
    ids = []
    for text, d, seed, n in jobs:
        try:
            ids.append(srv.submit(text, d, seed, n))
        except tild.TildError as e:
            assert e.status == 7   # a job larger than the whole budget is refused at submission
    srv.drain(timeout=120)
    st = srv.stats()
    assert st["peak_bytes"] <= st["budget"]
    assert st["done"] == len(ids) and st["failed"] == 0
    srv.close()


def test_job_over_budget_refused(tmp_path):
    text, d = decks.melt3d(tmp_path / "a")
    srv = tild.Server(device=0, hbm_budget=1 << 16)
    with pytest.raises(tild.TildError) as ei:
        srv.submit(text, d, 1, 10)
    assert ei.value.status == 7
    with pytest.raises(tild.TildError) as ei:
        srv.submit(text, d, 1, 0)
    assert ei.value.status == 6
    srv.close()
