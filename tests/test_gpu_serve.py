"""Serving (vd_server_*): continuous batching of tessellation and Lloyd jobs.

* a job's results are bit-identical alone and in a mix with other jobs that arrive while it runs,
  with the jobs in the small-job lane, in the large-job lane (worker thread) or split between them;
* a small job submitted while a large one runs finishes first and equals the oracle;
* a Lloyd job equals the oracle's calculate/relax loop (tessellation.rs:482-495) bit for bit;
* admission respects the HBM budget; a job larger than the whole budget is refused.
"""
import numpy as np
import pytest

from conftest import bits
from oracle import OracleTess
from vorothree_amd import BoundingBox, Server, SphereGeometry, Wall
from vorothree_amd import _lib
from vorothree_amd import workloads as W

pytestmark = pytest.mark.gpu

UNIT = BoundingBox((0.0, 0.0, 0.0), (1.0, 1.0, 1.0))


def job_points(seed, n, clustered=False):
    g = W.PCG32(seed)
    if clustered:
        return W.clustered_points(g, n, 8, 0.05, 1.0)
    return g.uniform(n, (0, 0, 0), (1, 1, 1))


def run_until_done(s, jobs, max_steps=500):
    for _ in range(max_steps):
        if all(s.info(j)["state"] == Server.FINISHED for j in jobs):
            return
        s.step()
    raise AssertionError("jobs did not finish")


@pytest.mark.parametrize("small_job", [None, 4000, 0])    # all small / split lanes / all large
def test_job_alone_equals_in_mix(small_job):
    pts = job_points(101, 6000)
    sphere = [Wall(-1000, SphereGeometry((0.5, 0.5, 0.5), 0.45))]
    alone = Server(hbm_budget=1 << 30)
    ja = alone.submit(UNIT, (14, 14, 14), sphere, pts, iterations=3)
    run_until_done(alone, [ja])
    va, ca, ga = alone.result(ja)

    mix = Server(hbm_budget=1 << 30, small_job=small_job)
    others = [mix.submit(UNIT, None, [], job_points(200 + k, 3000 + 1500 * k, clustered=k % 2 == 1),
                         iterations=1 + (k % 3)) for k in range(3)]
    mix.step()                                   # the others are already running ...
    j = mix.submit(UNIT, (14, 14, 14), sphere, pts, iterations=3)   # ... when this one arrives
    others += [mix.submit(UNIT, None, [], job_points(300 + k, 2000 + 700 * k), iterations=2) for k in range(2)]
    run_until_done(mix, others + [j])
    vm, cm, gm = mix.result(j)
    assert np.array_equal(bits(va), bits(vm))
    assert np.array_equal(bits(ca), bits(cm))
    assert np.array_equal(bits(ga), bits(gm))
    assert mix.stats()["steps"] >= 3


def test_lloyd_job_matches_oracle():
    pts = job_points(111, 5000)
    s = Server(hbm_budget=1 << 30)
    j = s.submit(UNIT, (13, 13, 13), [], pts, iterations=5)
    run_until_done(s, [j])
    vol, cen, gen = s.result(j)
    o = OracleTess((0, 0, 0), (1, 1, 1), 13, 13, 13)
    o.set_generators(pts)
    o.calculate(threads=8)
    for _ in range(4):
        o.relax()
        o.calculate(threads=8)
    ov, oc, _, _ = o.reduced()
    assert np.array_equal(bits(gen.reshape(-1)), bits(o.generators()))
    assert np.array_equal(bits(vol), bits(ov))
    assert np.array_equal(bits(cen), bits(oc))
    info = s.info(j)
    assert info["runs"] == 5 and info["total_cells"] == 5 * 5000


def test_small_job_runs_beside_a_large_one():
    s = Server(hbm_budget=8 << 30, small_job=20_000)
    big = s.submit(UNIT, None, [], job_points(140, 400_000), iterations=6)
    s.step()                                     # the large job's first step runs on the worker
    small_pts = job_points(141, 5000)
    j = s.submit(UNIT, (12, 12, 12), [], small_pts)
    run_until_done(s, [big, j])
    assert s.info(j)["finish"] < s.info(big)["finish"]
    assert s.info(big)["runs"] == 6 and s.info(big)["total_cells"] == 6 * 400_000
    vol, cen, gen = s.result(j)
    o = OracleTess((0, 0, 0), (1, 1, 1), 12, 12, 12)
    o.set_generators(small_pts)
    o.calculate(threads=8)
    ov, oc, _, _ = o.reduced()
    assert np.array_equal(bits(vol), bits(ov))
    assert np.array_equal(bits(cen), bits(oc))
    assert np.array_equal(bits(gen.reshape(-1)), bits(o.generators()))


def job_charge(grid, pts, iterations=2):
    """What the server charges a job (the budget bytes it holds while the job runs alone)."""
    s = Server(hbm_budget=8 << 30)
    j = s.submit(UNIT, grid, [], pts, iterations=iterations)
    s.step()
    charge = s.pending()["bytes"]
    run_until_done(s, [j])
    return charge, s.memory()["reserve"]


def test_budget_admission_and_refusal():
    small = [job_points(120 + k, 4000) for k in range(3)]
    charge, reserve = job_charge((12, 12, 12), small[0])
    s = Server(hbm_budget=reserve + charge + charge // 2)   # one such job at a time
    jobs = [s.submit(UNIT, (12, 12, 12), [], p, iterations=2) for p in small]
    s.step()
    p = s.pending()
    assert p["running"] == 1 and p["queued"] == 2
    run_until_done(s, jobs)
    for j in jobs:
        assert s.info(j)["state"] == Server.FINISHED
    assert s.pending()["bytes"] == 0
    assert s.memory()["peak"] <= s.memory()["budget"]
    with pytest.raises(_lib.VdError) as e:
        s.submit(UNIT, None, [], job_points(130, 200_000))
    assert e.value.code == _lib.VD_ENOMEM
    # a refused job never blocks the queue
    j = s.submit(UNIT, (10, 10, 10), [], small[0])
    run_until_done(s, [j])
    for j0 in jobs:
        s.release(j0)
    # a budget that cannot hold the server's fixed reserve is refused
    with pytest.raises(_lib.VdError) as e:
        Server(hbm_budget=reserve)
    assert e.value.code == _lib.VD_EINVAL


def test_peak_memory_within_budget_on_clustered_lloyd_job():
    # 300k points in 8 clusters (sigma 0.05), 6 Lloyd rounds on the density-rule grid: the first
    # calculate's cells next to the voids grow the shell prefix. The budget holds the reserve and
    # the job's charge with 10 % to spare; the device's free memory, sampled by a second thread
    # during the steps, never drops by more than the budget, and the server's own count agrees.
    import threading
    from vorothree_amd import _hip
    pts = job_points(140, 300_000, clustered=True)
    charge, reserve = job_charge(None, pts, iterations=6)
    budget = reserve + charge + charge // 10
    s = Server(hbm_budget=budget)
    # the same kind of job, smaller, first: module loads, the runtime's scratch for the tiers the
    # clusters reach, and its pools are then in place before the measured run
    warm = s.submit(UNIT, None, [], job_points(141, 20_000, clustered=True), iterations=2)
    run_until_done(s, [warm])
    s.release(warm)
    _hip.synchronize()
    free0 = _hip.mem_free()
    held0 = s.memory()["held"]
    low = [free0]
    stop = threading.Event()

    def sample():
        while not stop.is_set():
            low[0] = min(low[0], _hip.mem_free())

    th = threading.Thread(target=sample)
    th.start()
    try:
        j = s.submit(UNIT, None, [], pts, iterations=6)
        run_until_done(s, [j], max_steps=200)
    finally:
        stop.set()
        th.join()
    m = s.memory()
    assert s.info(j)["state"] == Server.FINISHED and s.info(j)["runs"] == 6
    assert m["peak"] <= budget
    assert held0 + (free0 - low[0]) <= budget
