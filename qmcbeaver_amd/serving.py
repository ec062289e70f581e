"""Many molecules on one GPU: batched uneven DMC and a continuous-batching
serving loop.

Each molecule (job) owns one context: its walker ensemble, Slater inverses
and stored orbital rows stay resident in HBM for the job's lifetime, and its
work runs on the context's own HIP stream.  Host threads drive the contexts
(the library calls release the interpreter lock), so kernels of different
jobs overlap on the device while each job's arithmetic is exactly what it is
alone: the random streams are counter-based per (seed, step, walker) and no
reduction crosses contexts, so every job's per-block results are independent
of the mix.

* `BatchedDMC` advances a fixed set of molecules of uneven electron, basis
  and walker counts together.
* `Server` admits jobs from an arrival stream as worker slots and device
  memory free up (by the footprint `qmcb_plan_device_bytes` projects for the
  job), runs them in blocks of steps, and records each block's latency from
  the moment it became runnable (the job's arrival for its first block, the
  end of its previous block otherwise) to its completion.

Molecules follow `synth.make_system(-1, n_elec=...)`: nuclei drawn from
Z in {1, 6, 7, 8} until their charge matches the electron count, about 5
Cartesian s/p/d functions per electron, exponents log-uniform on
[0.05, 1e4].
"""
from __future__ import annotations

import math
import threading
import time
from collections import deque
from dataclasses import dataclass, field
from typing import Callable, Dict, List, Optional

import numpy as np

from . import synth
from ._abi import QMCB_DRIFT_UMRIGAR93_UNEQUAL
from .core import Context, plan_device_bytes

# device bytes a context holds beyond its arrays (stream, events, kernel
# arguments, allocator rounding), added to every projected footprint
CONTEXT_OVERHEAD = 8 << 20


# DMC weights stay 1 for the first 100 steps, as the reference's
# QMCWalker::reweight_walker does (QMCWalker.cpp:1075-1080)
REWEIGHT_AFTER = 100
# the reference's default drift and weights (QMCFlags.cpp:78-79):
# umrigar93_unequalelectrons drift, umrigar93 energy modification
DRIFT = QMCB_DRIFT_UMRIGAR93_UNEQUAL


def initial_trial_energy(res) -> float:
    """Trial energy of a fresh ensemble: the median local energy of the
    walkers that evaluated (a mean is dominated by the walkers that start
    next to a node)."""
    ok = res["singular"] == 0
    return float(np.median(res["e_local"][ok])) if ok.any() else 0.0


def job_molecule(n_elec: int, seed: int) -> synth.System:
    """Synthetic molecule with n_elec (even) electrons, B ~ 5 n_elec."""
    return synth.make_system(-1, seed=int(seed), n_elec=int(n_elec))


@dataclass
class Job:
    """One VMC or DMC request: `blocks` blocks of `block_steps` single-
    electron steps of `walkers` walkers of one molecule."""
    job_id: int
    n_elec: int
    walkers: int
    blocks: int
    kind: str = "vmc"
    block_steps: int = 100
    seed: int = 0
    arrival: float = 0.0  # seconds after the start of the trace
    # filled in by the run
    results: List[Dict] = field(default_factory=list)
    latencies_ms: List[float] = field(default_factory=list)
    moves: int = 0
    footprint: int = 0      # projected device bytes (admission)
    device_bytes: int = 0   # what the job's context holds after its last block
    admitted: Optional[float] = None
    finished: Optional[float] = None
    error: Optional[str] = None


def poisson_trace(n_jobs: int, rate: float, seed: int, dmc_fraction: float = 0.25,
                  e_range=(8, 256), moves_range=(1e3, 1e5), block_steps: int = 100,
                  max_blocks: int = 4) -> List[Job]:
    """Jobs arriving as a Poisson stream of `rate` per second (rate = inf:
    the same jobs all at t = 0).  Electron counts are even and uniform on
    e_range; the requested walker-moves M are log-uniform on moves_range; a
    job runs 1..max_blocks blocks of block_steps steps and
    ceil(M / (block_steps * blocks)) walkers."""
    rng = np.random.default_rng(seed)
    t = 0.0
    jobs = []
    for j in range(n_jobs):
        t += rng.exponential(1.0 / rate)
        ne = 2 * int(rng.integers(e_range[0] // 2, e_range[1] // 2 + 1))
        M = math.exp(rng.uniform(math.log(moves_range[0]), math.log(moves_range[1])))
        blocks = int(rng.integers(1, max_blocks + 1))
        W = max(1, int(math.ceil(M / (block_steps * blocks))))
        kind = "dmc" if rng.random() < dmc_fraction else "vmc"
        if kind == "dmc":
            W = max(W, 16)  # a DMC population needs room to branch
        jobs.append(Job(job_id=j, n_elec=ne, walkers=W, blocks=blocks, kind=kind,
                        block_steps=block_steps, seed=int(rng.integers(1, 2**31)),
                        arrival=t))
    return jobs


class JobRunner:
    """Runs one job's blocks on a context of its own (the device path)."""

    def __init__(self, precision: str = "fp32", tau_vmc: float = 0.1, tau_dmc: float = 0.01,
                 refresh: int = 256, basis_scratch_bytes: int = 256 << 20):
        self.precision = precision
        self.tau_vmc, self.tau_dmc = tau_vmc, tau_dmc
        self.refresh = refresh if precision != "fp64" else 0
        self.scratch = basis_scratch_bytes
        self._mols: Dict = {}
        self._lock = threading.Lock()

    def molecule(self, job: Job) -> synth.System:
        key = (job.n_elec, job.seed)
        with self._lock:
            if key not in self._mols:
                self._mols[key] = job_molecule(job.n_elec, job.seed)
            return self._mols[key]

    def capacity(self, job: Job) -> int:
        """Walker slots of the job's context (DMC keeps room to branch)."""
        return job.walkers if job.kind == "vmc" else 2 * job.walkers

    def footprint(self, job: Job) -> int:
        return plan_device_bytes(self.molecule(job).pack(), precision=self.precision,
                                 max_walkers=self.capacity(job),
                                 basis_scratch_bytes=self.scratch) + CONTEXT_OVERHEAD

    def __call__(self, job: Job, on_block: Callable[[Dict, int], None]) -> None:
        s = self.molecule(job)
        W = job.walkers
        with Context(s.pack(), precision=self.precision, max_walkers=self.capacity(job),
                     refresh_interval=self.refresh, seed=job.seed,
                     basis_scratch_bytes=self.scratch, drift_mod=DRIFT) as ctx:
            res = ctx.evaluate(np.arange(W), synth.initial_walkers(s, W, job.seed))
            step = 0
            if job.kind == "dmc":
                e_t = initial_trial_energy(res)
                ctx.dmc_init(W)
            for _ in range(job.blocks):
                if job.kind == "vmc":
                    st = ctx.vmc_steps(W, step, job.block_steps, self.tau_vmc)
                    row = {"sum_e_local": st.sum_e_local, "sum_accept": st.sum_accept,
                           "n_samples": int(st.n_samples)}
                    moves = W * job.block_steps
                else:
                    st = ctx.dmc_steps(step, job.block_steps, self.tau_dmc, e_t,
                                       target=W, reweight_after=REWEIGHT_AFTER,
                                       energy_modification=True)
                    e_t = st.e_trial
                    row = {"sum_w_e_local": st.sum_w_e_local, "sum_weights": st.sum_weights,
                           "population": int(st.population), "e_trial": st.e_trial}
                    moves = int(st.n_moves)
                step += job.block_steps
                job.device_bytes = ctx.device_bytes()
                on_block(row, moves)


class Server:
    """Continuous batching of jobs on one GPU.

    Jobs are admitted in arrival order as soon as a worker slot is free and
    their projected device footprint fits in `capacity_bytes` next to the
    jobs already running (a job that could never fit is rejected).  Each
    admitted job runs on its own worker thread and context; the kernels of
    all running jobs overlap on the device.  `runner` executes one job (the
    device path by default)."""

    def __init__(self, capacity_bytes: int, workers: int = 8,
                 runner: Optional[JobRunner] = None):
        self.capacity = int(capacity_bytes)
        self.workers = int(workers)
        self.runner = runner if runner is not None else JobRunner()
        self.cv = threading.Condition()
        self.pending: deque = deque()
        self.in_use = 0
        self.active = 0
        self.done: List[Job] = []
        self.rejected: List[Job] = []
        self.t0 = 0.0
        self.peak_bytes = 0
        self.peak_active = 0

    def _now(self) -> float:
        return time.perf_counter() - self.t0

    def _admit_next(self) -> Optional[Job]:
        """Head-of-line admission (called with the lock held)."""
        while self.pending:
            job = self.pending[0]
            if job.footprint > self.capacity:
                self.pending.popleft()
                job.error = "footprint exceeds the device capacity"
                self.rejected.append(job)
                continue
            if self.active < self.workers and self.in_use + job.footprint <= self.capacity:
                self.pending.popleft()
                self.in_use += job.footprint
                self.active += 1
                self.peak_bytes = max(self.peak_bytes, self.in_use)
                self.peak_active = max(self.peak_active, self.active)
                job.admitted = self._now()
                return job
            return None
        return None

    def _worker(self, closing: threading.Event):
        while True:
            with self.cv:
                job = self._admit_next()
                while job is None:
                    if closing.is_set() and not self.pending:
                        self.cv.notify_all()
                        return
                    self.cv.wait(0.005)
                    job = self._admit_next()
            ready = [job.arrival]

            def on_block(row, moves, job=job, ready=ready):
                now = self._now()
                job.latencies_ms.append(1e3 * (now - ready[0]))
                job.results.append(row)
                job.moves += moves
                ready[0] = now
            try:
                self.runner(job, on_block)
            except Exception as e:  # the job fails, the server goes on
                job.error = f"{type(e).__name__}: {e}"
            with self.cv:
                job.finished = self._now()
                self.in_use -= job.footprint
                self.active -= 1
                self.done.append(job)
                self.cv.notify_all()

    def run(self, jobs: List[Job]) -> Dict:
        """Replay an arrival trace (job.arrival in seconds from now) and
        serve it to completion; returns the run's summary."""
        for j in jobs:
            j.footprint = self.runner.footprint(j)
        closing = threading.Event()
        threads = [threading.Thread(target=self._worker, args=(closing,), daemon=True)
                   for _ in range(self.workers)]
        self.t0 = time.perf_counter()
        for t in threads:
            t.start()
        for j in sorted(jobs, key=lambda j: j.arrival):
            delay = j.arrival - self._now()
            if delay > 0:
                time.sleep(delay)
            with self.cv:
                self.pending.append(j)
                self.cv.notify_all()
        closing.set()
        with self.cv:
            self.cv.notify_all()
        for t in threads:
            t.join()
        wall = self._now()
        ok = [j for j in self.done if j.error is None]
        lat = np.array([x for j in ok for x in j.latencies_ms] or [0.0])
        moves = sum(j.moves for j in ok)
        return {"jobs": len(jobs), "completed": len(ok),
                "failed": [(j.job_id, j.error) for j in self.done if j.error is not None],
                "rejected": [j.job_id for j in self.rejected],
                "wall_s": wall, "walker_moves": moves, "walker_moves_per_s": moves / wall,
                "block_latency_ms": {"p50": float(np.percentile(lat, 50)),
                                     "p99": float(np.percentile(lat, 99)),
                                     "blocks": int(sum(len(j.latencies_ms) for j in ok))},
                "peak_bytes": self.peak_bytes, "peak_active": self.peak_active}


def run_alone(jobs: List[Job], runner: JobRunner) -> Dict[int, List[Dict]]:
    """Each job run by itself, one after another: the per-block results a
    job produces with the device to itself."""
    out = {}
    for j in jobs:
        rows: List[Dict] = []
        runner(j, lambda row, moves, rows=rows: rows.append(row))
        out[j.job_id] = rows
    return out


# ----------------------------------------------------------------------------
# batched uneven DMC
# ----------------------------------------------------------------------------

def batched_composition(n_mol: int, seed: int, e_range=(8, 200), w_range=(256, 4096)):
    """(electrons, walkers, molecule seed) of each molecule: electron counts
    even and uniform on e_range, walker populations uniform on w_range."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n_mol):
        ne = 2 * int(rng.integers(e_range[0] // 2, e_range[1] // 2 + 1))
        W = int(rng.integers(w_range[0], w_range[1] + 1))
        out.append((ne, W, int(rng.integers(1, 2**31))))
    return out


class BatchedDMC:
    """A fixed set of molecules advanced together with DMC (tau, branching
    on FP64 weights), one context and host thread per molecule."""

    def __init__(self, composition, precision: str = "fp32", tau: float = 0.01,
                 workers: int = 16, basis_scratch_bytes: int = 256 << 20,
                 replace_cusps: bool = False):
        self.comp = list(composition)
        self.tau = tau
        self.workers = max(1, min(int(workers), len(self.comp)))
        self.precision = precision
        self.scratch = basis_scratch_bytes
        self.cusps = bool(replace_cusps)
        self.ctx: List[Optional[Context]] = [None] * len(self.comp)
        self.state = [dict(step=0, e_t=0.0, moves=0, pop=0) for _ in self.comp]

    def footprint(self) -> int:
        return sum(plan_device_bytes(job_molecule(ne, sd).pack(), precision=self.precision,
                                     max_walkers=2 * W, basis_scratch_bytes=self.scratch,
                                     replace_cusps=self.cusps)
                   + CONTEXT_OVERHEAD for ne, W, sd in self.comp)

    def _pool(self, fn, items):
        errors: List[str] = []
        lock = threading.Lock()
        it = iter(list(items))

        def body():
            while True:
                with lock:
                    i = next(it, None)
                if i is None:
                    return
                try:
                    fn(i)
                except Exception as e:
                    with lock:
                        errors.append(f"molecule {i}: {type(e).__name__}: {e}")
        th = [threading.Thread(target=body) for _ in range(self.workers)]
        for t in th:
            t.start()
        for t in th:
            t.join()
        if errors:
            raise RuntimeError("; ".join(errors[:3]))

    def setup(self):
        def one(i):
            ne, W, sd = self.comp[i]
            s = job_molecule(ne, sd)
            c = Context(s.pack(), precision=self.precision, max_walkers=2 * W,
                        refresh_interval=256 if self.precision != "fp64" else 0, seed=sd,
                        basis_scratch_bytes=self.scratch, drift_mod=DRIFT,
                        replace_cusps=self.cusps)
            self.ctx[i] = c
            res = c.evaluate(np.arange(W), synth.initial_walkers(s, W, sd))
            self.state[i]["e_t"] = initial_trial_energy(res)
            c.dmc_init(W)
        self._pool(one, range(len(self.comp)))

    def advance(self, steps: int, rows: Optional[List] = None, stats: Optional[List] = None):
        """`steps` single-electron DMC steps of every molecule; rows[i] gets
        molecule i's (sum w E_L, sum w, population, E_T), stats[i] its
        branching counts of these steps."""
        def one(i):
            st_i = self.state[i]
            W = self.comp[i][1]
            st = self.ctx[i].dmc_steps(st_i["step"], steps, self.tau, st_i["e_t"], target=W,
                                       reweight_after=REWEIGHT_AFTER, energy_modification=True)
            st_i["e_t"] = st.e_trial
            st_i["step"] += steps
            st_i["moves"] += int(st.n_moves)
            st_i["pop"] = int(st.population)
            if rows is not None:
                rows[i] = (st.sum_w_e_local, st.sum_weights, int(st.population), st.e_trial)
            if stats is not None:
                stats[i] = {k: getattr(st, k) for k in (
                    "n_split", "n_killed", "n_fused", "n_deleted", "n_split_skipped",
                    "population", "total_weight", "e_trial", "e_estimated", "sum_accept",
                    "n_moves")}
        self._pool(one, range(len(self.comp)))

    def close(self):
        for c in self.ctx:
            if c is not None:
                c.close()
        self.ctx = [None] * len(self.comp)

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def run(self, steps: int, warmup: int) -> Dict:
        t0 = time.perf_counter()
        self.setup()
        t_setup = time.perf_counter() - t0
        self.advance(warmup)
        for s in self.state:
            s["moves"] = 0
        stats: List = [None] * len(self.comp)
        t0 = time.perf_counter()
        self.advance(steps, stats=stats)
        wall = time.perf_counter() - t0
        moves = sum(s["moves"] for s in self.state)
        branching = {k: int(sum(s[k] for s in stats))
                     for k in ("n_split", "n_fused", "n_killed", "n_deleted", "n_split_skipped")}
        return {"molecules": len(self.comp), "steps": steps, "warmup": warmup,
                "replace_cusps": self.cusps,
                "electrons": [c[0] for c in self.comp],
                "walkers": [c[1] for c in self.comp],
                "population_end": [s["pop"] for s in self.state],
                "branching_in_timed_steps": branching,
                "wall_s": wall, "setup_s": t_setup, "walker_moves": moves,
                "walker_moves_per_s": moves / wall}
