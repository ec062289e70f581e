"""Host logic of the many-molecule paths, no device needed.

* qmcb_plan_device_bytes (host-only) against a recount of the arrays from
  their shapes for a plain single-determinant system, and its growth in the
  walker count, precision and determinant tables;
* the Server's admission: jobs start in arrival order, never more running
  than worker slots, never more projected bytes than the capacity, a job
  larger than the capacity rejected without stopping the others, a failing
  job reported without stopping the others, and each block's latency taken
  from the moment it became runnable;
* the Poisson trace's shape (rate, size ranges, requested walker-moves).
"""
import threading
import time

import numpy as np
import pytest

from qmcbeaver_amd import serving, synth
from qmcbeaver_amd.core import plan_device_bytes


def test_plan_grows_with_walkers_and_precision():
    s = synth.make_system(-1, seed=5, n_elec=40)
    p = s.pack()
    b = {prec: plan_device_bytes(p, precision=prec, max_walkers=512)
         for prec in ("bf16", "fp32", "fp32store", "fp64")}
    assert b["bf16"] < b["fp32"] < b["fp32store"] < b["fp64"]
    small = plan_device_bytes(p, precision="fp32", max_walkers=256)
    big = plan_device_bytes(p, precision="fp32", max_walkers=4096)
    # per-walker resident state: at least inverse + stored orbital rows
    n = 20
    per_walker = 2 * (8 * n * n + 8 * n * n + 4 * 4 * n * n)
    assert big - small >= (4096 - 256) * per_walker
    # a scratch cap below the point count bounds the basis scratch
    capped = plan_device_bytes(p, precision="fp32", max_walkers=4096,
                               basis_scratch_bytes=16 << 20)
    assert capped < big
    assert plan_device_bytes(p, precision="fp32", max_walkers=4096, replace_cusps=True) > big


def test_plan_resident_state_recount():
    """Difference of two plans that only differ in the walker count equals
    the per-walker arrays recounted from their shapes (no basis scratch
    change: the scratch cap is the same and W-bound on both sides)."""
    s = synth.make_system(-1, seed=9, n_elec=16)
    p = s.pack()
    n = [8, 8]
    N = 16
    B = int(p["n_basis"])
    kpad = -(-B // 32) * 32
    for prec in ("fp32", "fp64", "fp32store"):
        ts = 8 if prec == "fp64" else 4
        tx = 8 if prec != "fp32" else 4
        W1, W2 = 1024, 2048
        scratch = 64 * 5 * kpad * tx  # 64 points: the scratch is W-sized on both sides
        d = (plan_device_bytes(p, precision=prec, max_walkers=W2, basis_scratch_bytes=scratch)
             - plan_device_bytes(p, precision=prec, max_walkers=W1, basis_scratch_bytes=scratch))
        per = 0
        for k in n:
            per += 8 * k * k + 8 * k * k + 4 * k * k * ts  # inverse, values, derivatives
            per += 8 * 3 + 4 * 2                           # logdet sgn lap | singular dcount
            per += 8 * 3 * k                               # gradient ratios
            per += 2 * 8 * 8 * k                           # delayed-update buffers
        per += 8 * N * 3 * 6                               # R Ux Rin pUx grad drift
        per += 8 * 3 * 2                                   # propR gauss
        # per-walker doubles: U, Uxx Vee Ven cur_el, unif, pU, pUxx pVee pVen,
        # 4 energy accumulators, 3 observables, weight acc_w, ln|Psi|, 8 outputs,
        # mod_ratio, r2
        per += 8 * (1 + 4 + 1 + 1 + 3 + 4 + 3 + 2 + 1 + 8 + 1 + 3)
        per += 4 * (1 + 2 + 1)                             # active, pairs, ids
        per += 4 * 2                                       # age, refresh map
        per += 1 + 1 + 1                                   # mask, output sign, output singular flag
        per += 5 * 16 * tx                                 # orbital rows (64-aligned; 16 padded orbitals)
        per += 5 * kpad * tx                               # basis scratch rows (W points)
        per += 8 * 8 * 8 + 8 * 8 * 3 + 8 * 3 + 4 * 1 + 4  # staging: inverse, grad, lap/logdet/sgn, singular, row flag
        assert d == (W2 - W1) * per, (prec, d, (W2 - W1) * per)


class FakeRunner:
    """Stands in for the device: sleeps per block, records concurrency."""

    def __init__(self, block_s=0.02, fail=()):
        self.block_s = block_s
        self.fail = set(fail)
        self.lock = threading.Lock()
        self.running = 0
        self.max_running = 0
        self.order = []

    def footprint(self, job):
        return job.walkers * 1000

    def __call__(self, job, on_block):
        with self.lock:
            self.running += 1
            self.max_running = max(self.max_running, self.running)
            self.order.append(job.job_id)
        try:
            if job.job_id in self.fail:
                raise RuntimeError("planted failure")
            for b in range(job.blocks):
                time.sleep(self.block_s)
                on_block({"block": b, "job": job.job_id}, job.walkers * job.block_steps)
        finally:
            with self.lock:
                self.running -= 1


def test_server_admission_capacity_and_order():
    jobs = [serving.Job(job_id=i, n_elec=8, walkers=w, blocks=2, arrival=0.001 * i)
            for i, w in enumerate([30, 30, 30, 50, 20, 200, 10, 40])]
    r = FakeRunner(fail={6})
    srv = serving.Server(capacity_bytes=100_000, workers=3, runner=r)
    seen_bytes = []
    orig = srv._admit_next

    def watched():
        j = orig()
        seen_bytes.append(srv.in_use)
        return j
    srv._admit_next = watched
    out = srv.run(jobs)
    assert out["rejected"] == [5]            # 200 walkers * 1000 B > capacity
    assert [f[0] for f in out["failed"]] == [6]
    assert out["completed"] == 6
    assert max(seen_bytes) <= 100_000
    assert srv.peak_active <= 3 and r.max_running >= 2
    assert r.order == [0, 1, 2, 3, 4, 6, 7]  # arrival order, head of line
    for j in srv.done:
        if j.error is None:
            assert len(j.latencies_ms) == 2 and all(x > 0 for x in j.latencies_ms)
            assert j.latencies_ms[0] >= 1e3 * (j.admitted - j.arrival) + 15
    assert out["walker_moves"] == sum(j.walkers * 200 for j in jobs if j.job_id not in (5, 6))
    assert out["block_latency_ms"]["blocks"] == 12


def test_poisson_trace_shape():
    jobs = serving.poisson_trace(2000, 20.0, seed=3)
    t = np.array([j.arrival for j in jobs])
    gaps = np.diff(np.concatenate([[0.0], t]))
    assert abs(gaps.mean() - 0.05) < 0.005
    ne = np.array([j.n_elec for j in jobs])
    assert ne.min() >= 8 and ne.max() <= 256 and np.all(ne % 2 == 0)
    M = np.array([j.walkers * j.blocks * j.block_steps for j in jobs])
    vmc = np.array([j.kind == "vmc" for j in jobs])
    assert M[vmc].min() >= 1e3 and M[vmc].max() <= 1e5 * 1.05
    assert 0.15 < 1 - vmc.mean() < 0.35


def test_batched_composition_ranges():
    comp = serving.batched_composition(64, seed=3)
    assert len(comp) == 64
    assert all(8 <= ne <= 200 and ne % 2 == 0 and 256 <= W <= 4096 for ne, W, _ in comp)
    s = serving.job_molecule(comp[0][0], comp[0][2])
    assert abs(s.n_basis - 5 * s.n_elec) <= 0.2 * 5 * s.n_elec + 12
