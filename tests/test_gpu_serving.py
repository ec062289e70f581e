"""Many molecules on one device: footprint projection, isolation of jobs
in the mix, and the serving loop's admission with real contexts.

* qmcb_plan_device_bytes equals the arena a context allocates at creation,
  which holds every array including the staging of the first single-electron
  evaluate (single determinant, every precision, several determinants), and
  the context holds 32 bytes more per recorded sweep of the series (with
  cusp replacement the arena reserves the cusp tables at their maximum);
* isolation: every job's per-block results (VMC energy and acceptance sums,
  DMC weighted energy sums, populations, trial energies) are bit-identical
  when the job runs alone and when it runs in the mix with others on the
  device at the same time -- the random streams are counter-based and no
  reduction crosses contexts, so the stated tolerance is exact equality;
* batched DMC: the same molecules advanced by 1 host thread and by 6 give
  identical per-molecule results;
* the serving loop with a capacity that forces queueing never holds more
  projected bytes than the capacity, and every job's context stayed within
  its projection;
* at the two many-molecule workloads' full size: the 64-molecule batched
  DMC composition (150 243 walkers, 8-198 electrons, the reference's default
  cusp replacement, drift and energy modification) run past the unit-weight
  phase (steps 100-499 reweight and branch) keeps every population alive
  while splitting, fusing and deleting walkers, and three of its molecules
  run alone give the same per-molecule results bit for bit; the 200-job
  Poisson trace at 20 jobs/s (8-256 electrons, 1e3-1e5 walker-moves per job)
  completes, and its largest jobs re-run alone give the same per-block
  results bit for bit.
"""
import os
import tempfile
import time
import gzip

import numpy as np
import pytest

from qmcbeaver_amd import Context, ckmf, serving, synth
from qmcbeaver_amd.core import plan_device_bytes
from tests.inputs_util import GOLDEN

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def device_ready():
    """One small context from the main thread before the worker-thread tests,
    so the device runtime is initialised whatever order the files run in."""
    s = synth.make_system(0)
    with Context(s.pack(), precision="fp64", max_walkers=1) as c:
        c.evaluate(np.arange(1), synth.initial_walkers(s, 1, 1))


def _mcscf(name):
    with tempfile.TemporaryDirectory() as tmp:
        raw = os.path.join(tmp, name + ".ckmf")
        with gzip.open(os.path.join(GOLDEN, name + ".ckmf.gz"), "rb") as g, open(raw, "wb") as f:
            f.write(g.read())
        packed, _, _ = ckmf.read_ckmf(raw)
    return packed


@pytest.mark.parametrize("precision", ["fp32", "fp64", "fp32store", "bf16", "bf16store"])
@pytest.mark.parametrize("n_elec,W,scratch", [(16, 100, 0), (64, 300, 1 << 20)])
def test_plan_equals_allocation(precision, n_elec, W, scratch):
    p = synth.make_system(-1, seed=11 + n_elec, n_elec=n_elec).pack()
    plan = plan_device_bytes(p, precision=precision, max_walkers=W, basis_scratch_bytes=scratch)
    s = synth.make_system(-1, seed=11 + n_elec, n_elec=n_elec)
    R = synth.initial_walkers(s, W, 3)
    with Context(p, precision=precision, max_walkers=W, basis_scratch_bytes=scratch) as c:
        assert c.device_bytes() == plan  # the arena, allocated at creation
        c.evaluate(np.arange(W), R)
        R[:, 0] += 0.1
        c.evaluate(np.arange(W), R, whichE=0)  # staging arrays come out of the arena
        assert c.device_bytes() == plan
        c.vmc_steps(W, 0, 2, 0.1)  # + one row of the per-sweep series (32 bytes)
        assert c.device_bytes() == plan + 32


def test_plan_multidet_and_cusps():
    p = _mcscf("be_mcscf")
    W = 64
    R = np.random.default_rng(1).normal(size=(W, int(p["n_up"]) + int(p["n_dn"]), 3))
    plan = plan_device_bytes(p, precision="fp64", max_walkers=W)
    with Context(p, precision="fp64", max_walkers=W) as c:
        c.evaluate(np.arange(W), R)
        c.vmc_steps(W, 0, 2, 0.1)  # several determinants stage moves; one series row
        assert c.device_bytes() == plan + 32
    plan_c = plan_device_bytes(p, precision="fp64", max_walkers=W, replace_cusps=True)
    assert plan < plan_c
    with Context(p, precision="fp64", max_walkers=W, replace_cusps=True) as c:
        c.evaluate(np.arange(W), R)
        c.vmc_steps(W, 0, 2, 0.1)
        assert c.device_bytes() == plan_c + 32


def _small_trace():
    jobs = serving.poisson_trace(8, 200.0, seed=21, e_range=(8, 48), moves_range=(2e3, 2e4),
                                 block_steps=40, max_blocks=3, dmc_fraction=0.4)
    assert any(j.kind == "dmc" for j in jobs) and any(j.kind == "vmc" for j in jobs)
    return jobs


def _fresh(jobs):
    return [serving.Job(job_id=j.job_id, n_elec=j.n_elec, walkers=j.walkers, blocks=j.blocks,
                        kind=j.kind, block_steps=j.block_steps, seed=j.seed, arrival=j.arrival)
            for j in jobs]


@pytest.mark.parametrize("precision", ["fp32", "fp64"])
def test_job_results_same_alone_and_in_the_mix(precision):
    jobs = _small_trace()
    runner = serving.JobRunner(precision=precision)
    alone = serving.run_alone(_fresh(jobs), runner)
    mixed = _fresh(jobs)
    srv = serving.Server(capacity_bytes=64 << 30, workers=4, runner=runner)
    out = srv.run(mixed)
    assert out["completed"] == len(jobs), out["failed"]
    assert srv.peak_active >= 2  # the jobs did overlap
    for j in mixed:
        assert len(j.results) == j.blocks
        assert j.results == alone[j.job_id], j.job_id
        assert 0 < j.device_bytes <= j.footprint


def test_batched_dmc_thread_independent():
    comp = [(ne, W, sd) for ne, W, sd in serving.batched_composition(6, seed=5, e_range=(8, 40),
                                                                     w_range=(64, 256))]
    res = []
    for workers in (1, 6):
        rows = [None] * len(comp)
        with serving.BatchedDMC(comp, precision="fp32", workers=workers) as b:
            b.setup()
            b.advance(30)
            b.advance(30, rows)
        res.append(rows)
    assert res[0] == res[1]
    for sw_e, sw, pop, e_t in res[0]:
        assert sw > 0 and pop > 0 and np.isfinite(sw_e / sw) and np.isfinite(e_t)


def test_server_capacity_forces_queueing():
    jobs = _small_trace()
    runner = serving.JobRunner(precision="fp32")
    fps = sorted(runner.footprint(j) for j in jobs)
    cap = fps[-1] + fps[-2] // 2  # at most one large job at a time
    srv = serving.Server(capacity_bytes=cap, workers=4, runner=runner)
    out = srv.run(_fresh(jobs))
    assert out["completed"] == len(jobs), out["failed"]
    assert srv.peak_bytes <= cap
    assert out["block_latency_ms"]["p99"] >= out["block_latency_ms"]["p50"] > 0


CFG3_SEED = 0x51EED03


def test_batched_dmc_config3_full_composition_past_step_100():
    comp = serving.batched_composition(64, CFG3_SEED)
    assert len(comp) == 64 and sum(c[1] for c in comp) == 150243
    print("case batched DMC config 3: 64 molecules, cusp replacement, steps 0-499", flush=True)
    rows, stats = [None] * 64, [None] * 64
    with serving.BatchedDMC(comp, precision="fp32", workers=16, replace_cusps=True) as b:
        b.setup()
        b.advance(100)  # weights stay 1 for the first 100 steps
        t0 = time.perf_counter()
        b.advance(400, rows, stats)
        wall = time.perf_counter() - t0
    moves = sum(st["n_moves"] for st in stats)
    print("steps 100-499: %.3e walker-moves in %.2f s, %.3e walker-moves/s"
          % (moves, wall, moves / wall), flush=True)
    tot = {k: sum(st[k] for st in stats) for k in
           ("n_split", "n_fused", "n_killed", "n_deleted", "n_split_skipped")}
    W = np.array([c[1] for c in comp])
    pop = np.array([st["population"] for st in stats])
    branched = sum(1 for st in stats if st["n_split"] + st["n_fused"] + st["n_killed"] > 0)
    print("branching", tot, "molecules branched", branched, "population/target min %.2f max %.2f"
          % ((pop / W).min(), (pop / W).max()), flush=True)
    assert tot["n_split"] > 0 and tot["n_fused"] > 0 and tot["n_killed"] > 0
    assert branched == 64  # every molecule reweighted and branched after step 100
    assert np.all(pop >= 0.5 * W) and np.all(pop <= 2 * W)
    for sw_e, sw, p, e_t in rows:
        assert sw > 0 and np.isfinite(sw_e / sw) and np.isfinite(e_t)
    # isolation at full size: the smallest, the largest and the most populous
    # molecule run by themselves
    pick = sorted({0, int(np.argmax([c[0] for c in comp])), int(np.argmax(W))})
    alone = [None] * len(pick)
    with serving.BatchedDMC([comp[i] for i in pick], precision="fp32", workers=1,
                            replace_cusps=True) as b:
        b.setup()
        b.advance(100)
        b.advance(400, alone)
    for k, i in enumerate(pick):
        assert alone[k] == rows[i], i


def test_serving_config4_trace_and_isolation():
    jobs = serving.poisson_trace(200, 20.0, 20)  # the benchmark's trace
    runner = serving.JobRunner(precision="fp32")
    print("case serving config 4: 200 jobs at 20/s", flush=True)
    srv = serving.Server(capacity_bytes=160 << 30, workers=8, runner=runner)
    out = srv.run(jobs)
    print("completed", out["completed"], "wall", out["wall_s"], "latency",
          out["block_latency_ms"], "peak active", out["peak_active"], flush=True)
    assert out["completed"] == 200, out["failed"]
    assert out["rejected"] == [] and srv.peak_active >= 2
    assert max(j.n_elec for j in jobs) >= 240
    moves = {j.job_id: j.walkers * j.block_steps * j.blocks for j in jobs}
    sel = {max((j for j in jobs if j.kind == k), key=lambda j: moves[j.job_id]).job_id
           for k in ("vmc", "dmc")}
    sel.add(max(jobs, key=lambda j: j.n_elec).job_id)
    sel.add(max(jobs, key=lambda j: moves[j.job_id] * j.n_elec).job_id)
    chosen = [j for j in jobs if j.job_id in sel]
    print("re-run alone", [(j.job_id, j.kind, j.n_elec, j.walkers, j.blocks) for j in chosen],
          flush=True)
    alone = serving.run_alone(_fresh(chosen), runner)
    for j in chosen:
        assert len(j.results) == j.blocks
        assert j.results == alone[j.job_id], j.job_id
