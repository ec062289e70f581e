"""Walker-moves/s of the MI355X walker-evaluation path.

Workload (default): the 256-electron synthetic molecule (128 up + 128 down,
40 nuclei, 1200 Cartesian s/p/d basis functions), 8192 walkers resident in
HBM, single-electron importance-sampled VMC moves.  One step = every walker
proposes a move of electron (step mod 256), the full evaluate contract runs at
the proposal (ln|Psi|, sign, grad ln Psi of all electrons, Laplacian, E_L) and
the move is accepted or rejected on the device.  The mixed path (basis in
FP64 rounded to FP32 for the MFMA orbital GEMM, FP32 derivative matrices,
Sherman-Morrison inverse/log-det/E_L in FP64, FP64 re-inversion every 256
moves of a spin), with the walkers in 3 overlapping groups, is
`value`; the plain FP64 path on the same GPU, the fp32store and bf16 modes,
their per-quantity errors against the FP64 path, a serving run of 200
Poisson-arriving jobs and the reference's own CPU evaluate on the host cores
are reported beside it.

    python bench.py [--gpus 1] [--steps K] [--warmup W]
    python bench.py --workload batched     # uneven DMC over 64 molecules
    python bench.py --workload serving     # continuous-batching serving loop
"""
from __future__ import annotations

import argparse
import json
import os
import subprocess
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "walker-moves/s at 256 e⁻ (mixed vs FP64, same GPU) + p50/p99 block latency"
HBM_PEAK_GBS = 8000.0
PROFILE = os.path.join(ROOT, "profiles", "sm_move_mol256.md")


def measured_traffic(precision):
    """HBM bytes per sm_move launch of this workload, from the rocprofv3
    FETCH_SIZE / WRITE_SIZE passes summarised in profiles/sm_move_mol256.md
    (scaled by the plain-copy calibration stated there); None if absent."""
    try:
        for line in open(PROFILE):
            if line.startswith("traffic_bytes_per_launch:"):
                kv = dict(t.split("=") for t in line.split(":", 1)[1].split())
                return float(kv[{"fp64": "fp64", "bf16store": "bf16store"}.get(precision, "mixed")])
    except (OSError, KeyError, ValueError):
        pass
    return None


def sm_bytes_per_move(n, N, s_d, s_o, acc):
    """Algorithmic HBM bytes of one single-electron move in the sm_move kernel:
    read A^-1 (8n^2) and the four derivative matrices (4 s_d n^2), the new
    orbital rows (5 n s_o), current and proposed grad U (2 x 24N), both spins'
    grad D/D (24N); on acceptance write A^-1 (8n^2), the new column of the five
    stored matrices (5 n s_d), grad D/D of the moved spin (24n) and grad U (24N)."""
    reads = 8 * n * n + 4 * s_d * n * n + 5 * n * s_o + 72 * N
    writes = acc * (8 * n * n + 5 * n * s_d + 24 * n + 24 * N)
    return reads + writes


def run_path(s, precision, W, steps, warmup, tau, block, refresh, seed, groups=1):
    """Timed VMC of W walkers: `steps` steps in blocks of `block`, with the
    walkers split into `groups` overlapping groups (bit-identical results for
    every group count), then a second pass of the same length with one group
    and per-kernel HIP events on the library's stream (the kernel times and
    the roofline come from that pass; the events cost ~3 % and stay out of
    the timed throughput)."""
    from qmcbeaver_amd import synth, Context
    R0 = synth.initial_walkers(s, W, seed)
    ctx = Context(s.pack(), precision=precision, max_walkers=W, refresh_interval=refresh,
                  seed=seed, walker_groups=groups)
    ids = np.arange(W)
    t0 = time.perf_counter()
    res = ctx.evaluate(ids, R0)
    t_init = time.perf_counter() - t0
    n_sing = int(res["singular"].sum())
    step = 0
    if warmup:
        ctx.vmc_steps(W, step, warmup, tau)
        step += warmup
    ctx.sync()
    lat = []
    acc = 0.0
    esum = 0.0
    t0 = time.perf_counter()
    done = 0
    while done < steps:
        k = min(block, steps - done)
        tb = time.perf_counter()
        st = ctx.vmc_steps(W, step, k, tau)
        if k == block:  # latency statistics over full blocks only
            lat.append((time.perf_counter() - tb) * 1e3)
        acc += st.sum_accept
        esum += st.sum_e_local
        step += k
        done += k
    ctx.sync()
    dt = time.perf_counter() - t0
    groups_used = ctx.walker_groups()
    ctx.set_walker_groups(1)
    ctx.enable_timing(True)
    tk = time.perf_counter()
    ctx.vmc_steps(W, step, steps, tau)
    ctx.sync()
    dtk = time.perf_counter() - tk
    kt = ctx.kernel_times()
    ctx.enable_timing(False)
    ctx.close()
    return {
        "walker_moves_per_s": W * steps / dt,
        "ms_per_step": 1e3 * dt / steps,
        "ms_per_step_one_group_timed": 1e3 * dtk / steps,
        "block_ms": lat,
        "acceptance": acc / (W * steps),
        "mean_e_local": esum / (W * steps),
        "init_s": t_init,
        "init_singular": n_sing,
        "kernels": kt,
        "walker_groups": groups_used,
    }


def equilibrated_configs(s, n_conf=10000, seed=4321, tau=0.3, eq_sweeps=100, gap_sweeps=4,
                         walkers=2048):
    """n_conf configurations drawn from an equilibrated FP64 device VMC with
    the reference's default proposal (Umrigar-93 modified drift for unequal
    electrons, QMCFlags.cpp:77-78): `walkers` walkers from N(R_A, 0.5^2),
    eq_sweeps sweeps of single-electron moves at tau (the mean local energy
    and the acceptance are flat from about sweep 60 at tau = 0.3), then one
    snapshot of every walker each gap_sweeps sweeps until n_conf are taken."""
    from qmcbeaver_amd import synth, Context
    from qmcbeaver_amd._abi import QMCB_DRIFT_UMRIGAR93_UNEQUAL
    N = s.n_elec
    W = min(walkers, n_conf)
    ids = np.arange(W)
    snaps = []
    with Context(s.pack(), precision="fp64", max_walkers=W, seed=seed,
                 drift_mod=QMCB_DRIFT_UMRIGAR93_UNEQUAL, dt=tau) as c:
        c.evaluate(ids, synth.initial_walkers(s, W, seed))
        step = 0
        c.vmc_steps(W, step, eq_sweeps * N, tau)
        step += eq_sweeps * N
        while sum(len(x) for x in snaps) < n_conf:
            snaps.append(c.positions(W))
            c.vmc_steps(W, step, gap_sweeps * N, tau)
            step += gap_sweeps * N
    R = np.concatenate(snaps)[:n_conf]
    desc = (f"FP64 device VMC, umrigar93 modified drift, tau {tau}: {W} walkers after "
            f"{eq_sweeps} sweeps, {len(snaps)} snapshots {gap_sweeps} sweeps apart")
    return R, desc


def accuracy_report(s, R, desc, precision="fp32"):
    """Per-quantity error of a mixed mode against the FP64 path on the
    configurations R (equilibrated_configs), all-electron evaluates and a
    single-electron proposal (moved electron 0, Sherman-Morrison on the
    mode's stored rows); and against the values the reference program wrote
    for the four golden configurations of the 256-electron molecule.  The
    FP64 path matches the reference evaluate to <= 1e-10."""
    from qmcbeaver_amd import Context
    N = s.n_elec
    n_conf = len(R)
    ids = np.arange(n_conf)
    Rp = R.copy()
    Rp[:, 0, :] += 0.05
    with Context(s.pack(), precision="fp64", max_walkers=n_conf) as c64:
        a64 = c64.evaluate(ids, R)
        p64 = c64.evaluate(ids, Rp, whichE=0)
    with Context(s.pack(), precision=precision, max_walkers=n_conf) as c32:
        a32 = c32.evaluate(ids, R)
        p32 = c32.evaluate(ids, Rp, whichE=0)
    ok = ~(a64["singular"].astype(bool) | a32["singular"].astype(bool) |
           p64["singular"].astype(bool) | p32["singular"].astype(bool))

    def stats(x, y, per=1.0):
        d = np.abs(np.asarray(x)[ok] - np.asarray(y)[ok]).reshape(ok.sum(), -1)
        d = d.max(axis=1) / per
        return {"max": float(d.max()), "rms": float(np.sqrt(np.mean(d * d))),
                "p99": float(np.percentile(d, 99)), "median": float(np.median(d)),
                "above_1e-6": float(np.mean(d > 1e-6))}
    out = {"configurations": int(ok.sum()), "sampler": desc, "reference": "FP64 path",
           "precision": precision,
           "all_electron": {"log_psi": stats(a32["log_psi"], a64["log_psi"]),
                            "grad_lnpsi": stats(a32["grad_lnpsi"], a64["grad_lnpsi"]),
                            "e_local_per_electron": stats(a32["e_local"], a64["e_local"], N)},
           "single_electron": {"log_psi": stats(p32["log_psi"], p64["log_psi"]),
                               "grad_lnpsi": stats(p32["grad_lnpsi"], p64["grad_lnpsi"]),
                               "e_local_per_electron": stats(p32["e_local"], p64["e_local"], N)},
           "sign_mismatch": int(np.sum(a32["sign_psi"][ok] != a64["sign_psi"][ok]))}
    gpath = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tests", "golden",
                         "cfg2_mol256.npz")
    if os.path.exists(gpath):
        g = np.load(gpath)
        with Context(s.pack(), precision=precision, max_walkers=len(g["R"])) as c:
            r = c.evaluate(np.arange(len(g["R"])), g["R"])
        out["reference_goldens"] = {
            "configurations": int(len(g["R"])),
            "e_local_per_electron_max": float(np.max(np.abs(r["e_local"] - g["ev_e_local"])) / N),
            "log_psi_max": float(np.max(np.abs(r["log_psi"] - g["ev_log_psi"])))}
    return out


def cpu_model():
    """The host CPU's model name (as lscpu prints it, from /proc/cpuinfo)."""
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(s, tau, seed, target_s=10.0):
    """The reference's own evaluate (oracle/_ref, compiled from its sources)
    timed on this host's cores, two legs: one single-threaded reference
    process alone (1 core), then one process per core (at most 16, the host
    share of one GPU), each on its own walkers, all started together, the
    rate being the sum of their walker-moves over the slowest one's time.
    Without the reference build: the C oracle port on one core."""
    from qmcbeaver_amd import synth
    drv = os.path.join(ROOT, "oracle", "_ref", "qmcb_ref_driver")
    nw = 8
    if os.path.exists(drv):
        try:
            cores = max(1, min(16, len(os.sched_getaffinity(0))))
        except AttributeError:
            cores = 1
        with tempfile.TemporaryDirectory() as tmp:
            ck = os.path.join(tmp, "in.ckmf")
            synth.write_ckmf(s, ck, dt=tau)
            cfs = []
            for c in range(cores):
                cf = os.path.join(tmp, f"cfg{c}.bin")
                with open(cf, "wb") as f:
                    np.array([nw, s.n_elec], dtype=np.int32).tofile(f)
                    synth.initial_walkers(s, nw, seed + c).tofile(f)
                cfs.append(cf)
            # calibrate on a short run, then a bounded sample of ~target_s
            out = subprocess.run([drv, "vmc", ck, cfs[0], "16", str(tau), "7"], cwd=tmp,
                                 capture_output=True, text=True, check=True).stdout
            r = json.loads(out.strip().splitlines()[-1])
            rate = r["walker_moves"] / max(r["seconds"], 1e-9)
            moves = int(max(16, min(4096, target_s * rate / nw)))
            print(f"cpu baseline: 1 core, {nw} walkers x {moves} moves", file=sys.stderr,
                  flush=True)
            out = subprocess.run([drv, "vmc", ck, cfs[0], str(moves), str(tau), "7"], cwd=tmp,
                                 capture_output=True, text=True, check=True).stdout
            one = json.loads(out.strip().splitlines()[-1])
            print(f"cpu baseline: {cores} processes", file=sys.stderr, flush=True)
            procs = [subprocess.Popen([drv, "vmc", ck, cf, str(moves), str(tau), str(7 + c)],
                                      cwd=tmp, stdout=subprocess.PIPE, stderr=subprocess.DEVNULL,
                                      text=True) for c, cf in enumerate(cfs)]
            res = []
            for pr in procs:
                o, _ = pr.communicate()
                if pr.returncode != 0:
                    raise RuntimeError("reference driver failed")
                res.append(json.loads(o.strip().splitlines()[-1]))
        total = sum(r["walker_moves"] for r in res)
        slowest = max(r["seconds"] for r in res)
        return {"value": total / slowest, "unit": "walker-moves/s", "cores": cores,
                "kind": "reference", "cpu_model": cpu_model(), "nproc": os.cpu_count(),
                "one_core": {"value": one["walker_moves"] / one["seconds"], "cores": 1,
                             "sample": f"1 process x {nw} walkers x {moves} single-electron "
                                       f"VMC moves, {one['seconds']:.1f} s"},
                "sample": f"{cores} processes x {nw} walkers x {moves} single-electron VMC moves "
                          f"of the same molecule (reference QMCSCFJastrow::evaluate, "
                          f"re-evaluate on rejection), slowest {slowest:.1f} s"}
    R = synth.initial_walkers(s, nw, seed)
    from oracle.pyoracle import Oracle
    o = Oracle(s.pack())
    t, _ = o.vmc_time(R, 32, tau)
    rate = nw * 32 / t
    moves = int(max(16, min(4096, target_s * rate / nw)))
    t, _ = o.vmc_time(R, moves, tau)
    return {"value": nw * moves / t, "unit": "walker-moves/s", "cores": 1, "kind": "port",
            "cpu_model": cpu_model(), "nproc": os.cpu_count(),
            "sample": f"{nw} walkers x {moves} single-electron VMC moves (C oracle), {t:.1f} s"}


def serving_warmup(runner):
    """Untimed: one small VMC and one small DMC job for each Sherman-Morrison
    column tile (n = 8, 24, 50, 80, 120 electrons per spin), run alone, so
    that the timed trace starts with the device code loaded, as in a
    long-running server.  Returns the number of warm-up jobs."""
    from qmcbeaver_amd import serving
    jobs = []
    for i, ne in enumerate((16, 48, 100, 160, 240)):
        for kind in ("vmc", "dmc"):
            jobs.append(serving.Job(job_id=-1 - len(jobs), n_elec=ne, walkers=16, blocks=1,
                                    kind=kind, block_steps=10, seed=1000 + i))
    serving.run_alone(jobs, runner)
    return len(jobs)


def serving_run(precision, n_jobs, rate, seed, workers, capacity_gib):
    """Serving loop: Poisson arrivals of VMC/DMC jobs (8-256 electrons,
    1e3-1e5 walker-moves each, blocks of 100 steps), continuous batching
    with admission by projected device footprint, after untimed warm-up
    jobs."""
    from qmcbeaver_amd import serving
    jobs = serving.poisson_trace(n_jobs, rate, seed)
    runner = serving.JobRunner(precision=precision)
    n_warm = serving_warmup(runner)
    srv = serving.Server(capacity_bytes=int(capacity_gib * (1 << 30)), workers=workers,
                         runner=runner)
    print(f"serving {precision}: {n_jobs} jobs at {rate}/s", file=sys.stderr, flush=True)
    out = srv.run(jobs)
    out["warmup_jobs"] = n_warm
    print(f"serving {precision}: {out['completed']} completed in {out['wall_s']:.1f} s",
          file=sys.stderr, flush=True)
    out["failed"] = out["failed"][:5]
    out["precision"] = precision
    out["arrival_rate_per_s"] = rate
    out["mean_jobs_in_flight"] = float(sum(j.finished - j.admitted for j in srv.done
                                           if j.admitted is not None) / max(out["wall_s"], 1e-9))
    return out


def batched_run(precision, n_mol, steps, warmup, seed, workers, tau):
    """Batched uneven DMC: n_mol molecules (8-200 electrons, B ~ 5 N,
    256-4096 walkers each), DMC with tau and the reference's default cusp
    replacement, drift and energy modification, all molecules advanced
    together; the timed steps follow `warmup` steps (from 100 on, every
    sweep reweights and branches)."""
    from qmcbeaver_amd import serving
    comp = serving.batched_composition(n_mol, seed)
    with serving.BatchedDMC(comp, precision=precision, tau=tau, workers=workers,
                            replace_cusps=True) as b:
        fp = b.footprint()
        print(f"batched {precision}: {n_mol} molecules, {sum(c[1] for c in comp)} walkers, "
              f"projected {fp / 2**30:.1f} GiB", file=sys.stderr, flush=True)
        out = b.run(steps, warmup)
        print(f"batched {precision}: setup {out['setup_s']:.1f} s, timed {out['wall_s']:.2f} s",
              file=sys.stderr, flush=True)
    out["projected_device_bytes"] = fp
    out["precision"] = precision
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=256)
    ap.add_argument("--warmup", type=int, default=32)
    ap.add_argument("--walkers", type=int, default=8192)
    ap.add_argument("--config", type=int, default=2)
    ap.add_argument("--tau", type=float, default=0.01)
    ap.add_argument("--block", type=int, default=16, help="steps per timed block")
    ap.add_argument("--refresh", type=int, default=256,
                    help="FP64 re-inversion of each walker's matrices of a spin every this many "
                         "moves of that spin (mixed modes; the FP64 run never re-inverts)")
    ap.add_argument("--groups", type=int, default=3,
                    help="walker groups on overlapping streams (results are bit-identical); "
                         "-1: chosen by the library's timing")
    ap.add_argument("--precision", default="fp32",
                    choices=["fp32", "fp32store", "bf16", "bf16store", "fp64"])
    ap.add_argument("--no-fp64", action="store_true", help="skip the FP64 comparison run")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--accuracy", type=int, default=10000,
                    help="configurations for the mixed-vs-FP64 error report (0 = skip)")
    ap.add_argument("--workload", default="mol256", choices=["mol256", "batched", "serving"],
                    help="mol256: one 256-electron molecule (default); batched: uneven DMC "
                         "over many molecules; serving: Poisson arrival stream of jobs")
    ap.add_argument("--serving-jobs", type=int, default=200,
                    help="jobs in the serving trace (default 200 = 10 s of arrivals at 20/s; "
                         "0 = no serving run beside the mol256 workload)")
    ap.add_argument("--rate", type=float, default=20.0, help="serving arrivals per second")
    ap.add_argument("--serving-workers", type=int, default=8)
    ap.add_argument("--capacity-gib", type=float, default=160.0,
                    help="device memory the serving loop may hand to jobs")
    ap.add_argument("--molecules", type=int, default=64)
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a plain 1 GiB device-to-device copy three times (the "
                         "reference kernel for scaling the FETCH_SIZE / WRITE_SIZE counters)")
    args = ap.parse_args()
    if args.copy_calibration:
        import torch
        x = torch.ones(1 << 27, dtype=torch.float64, device="cuda")
        y = torch.empty_like(x)
        torch.cuda.synchronize()
        for _ in range(3):
            y.copy_(x)
        torch.cuda.synchronize()
        print(json.dumps({"copy_calibration_bytes": x.numel() * 8, "copies": 3}))
        return
    if args.gpus != 1:
        print(json.dumps({"error": "the walker-evaluation path is measured on one GPU only; "
                                   "it would shard by whole walkers, with no exchange during "
                                   "evaluate"}))
        sys.exit(2)

    if args.workload == "serving":
        mixed = serving_run(args.precision, args.serving_jobs, args.rate, 20, args.serving_workers,
                            args.capacity_gib)
        fp64 = None if args.no_fp64 else serving_run("fp64", args.serving_jobs, args.rate, 20,
                                                     args.serving_workers, args.capacity_gib)
        # the same jobs all arriving at once: the loop's capacity for this mix
        sat = serving_run(args.precision, args.serving_jobs, float("inf"), 20,
                          args.serving_workers, args.capacity_gib)
        print(json.dumps({
            "metric": "aggregate walker-moves/s + p50/p99 block latency (serving loop)",
            "value": mixed["walker_moves_per_s"], "unit": "walker-moves/s", "n_gpus": 1,
            "steps": args.serving_jobs, "warmup": mixed["warmup_jobs"],
            "ms_per_step": 1e3 * mixed["wall_s"] / max(args.serving_jobs, 1),
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
            "dtype": "f32+f64" if args.precision != "fp64" else "f64", "data": "synthetic",
            "config": {"workload": f"serving loop: {args.serving_jobs} VMC/DMC jobs, Poisson "
                                   f"arrivals {args.rate}/s, 8-256 e-, 1e3-1e5 walker-moves per "
                                   "job in blocks of 100 steps, admission by projected device "
                                   f"footprint ({args.capacity_gib} GiB), "
                                   f"{args.serving_workers} job slots"},
            "block_latency_ms": mixed["block_latency_ms"], "serving": mixed, "fp64": fp64,
            "saturated": {"walker_moves_per_s": sat["walker_moves_per_s"],
                          "wall_s": sat["wall_s"], "block_latency_ms": sat["block_latency_ms"],
                          "completed": sat["completed"],
                          "note": "the same trace with every job arriving at t = 0"}}))
        return
    if args.workload == "batched":
        mixed = batched_run(args.precision, args.molecules, args.steps, args.warmup, 0x51EED03,
                            16, 0.01)
        fp64 = None if args.no_fp64 else batched_run("fp64", args.molecules, args.steps,
                                                     args.warmup, 0x51EED03, 16, 0.01)
        print(json.dumps({
            "metric": "aggregate walker-moves/s (batched uneven DMC)",
            "value": mixed["walker_moves_per_s"], "unit": "walker-moves/s", "n_gpus": 1,
            "steps": args.steps, "warmup": args.warmup,
            "ms_per_step": 1e3 * mixed["wall_s"] / max(args.steps, 1),
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
            "dtype": "f32+f64" if args.precision != "fp64" else "f64", "data": "synthetic",
            "config": {"workload": f"batched uneven DMC: {args.molecules} molecules, 8-200 e-, "
                                   "B ~ 5 N, 256-4096 walkers each, tau=0.01, one step = one "
                                   "single-electron DMC step of every molecule"},
            "batched": mixed, "fp64": fp64,
            "mixed_over_fp64": None if fp64 is None else
            mixed["walker_moves_per_s"] / fp64["walker_moves_per_s"]}))
        return

    from qmcbeaver_amd import synth
    s = synth.make_system(args.config)
    n, N = s.n_up, s.n_elec
    mixed = run_path(s, args.precision, args.walkers, args.steps, args.warmup, args.tau,
                     args.block, args.refresh if args.precision != "fp64" else 0, 1234,
                     args.groups)
    fp64 = None
    store = None
    bf16 = None
    bf16store = None
    if not args.no_fp64 and args.precision != "fp64":
        fp64 = run_path(s, "fp64", args.walkers, args.steps, args.warmup, args.tau, args.block,
                        0, 1234, args.groups)
        if args.precision == "fp32":
            store = run_path(s, "fp32store", args.walkers, args.steps, args.warmup, args.tau,
                             args.block, args.refresh, 1234, args.groups)
            bf16 = run_path(s, "bf16", args.walkers, args.steps, args.warmup, args.tau,
                            args.block, args.refresh, 1234, args.groups)
            bf16store = run_path(s, "bf16store", args.walkers, args.steps, args.warmup,
                                 args.tau, args.block, args.refresh, 1234, args.groups)

    s_d = {"fp64": 8, "fp32store": 4, "bf16store": 2}.get(args.precision, 4)
    kt = mixed["kernels"].get("sm_move", {"ms": float("nan"), "launches": 1})
    avg_ms = kt["ms"] / max(kt["launches"], 1)
    bytes_launch = args.walkers * sm_bytes_per_move(n, N, s_d, s_d, mixed["acceptance"])
    achieved = bytes_launch / (avg_ms * 1e-3) / 1e9
    def block_stats(ms):
        """p50/p99 over the full blocks of the timed run (None without one)."""
        if not ms:
            return None
        return {"p50": float(np.percentile(ms, 50)), "p99": float(np.percentile(ms, 99)),
                "blocks": len(ms), "block_steps": args.block}
    out = {
        "metric": METRIC,
        "value": mixed["walker_moves_per_s"],
        "unit": "walker-moves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": mixed["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64" if args.precision != "fp64" else "f64",
        "data": "synthetic",
        "config": {
            "workload": f"{s.name}: {N} e- ({n}+{s.n_dn}), {s.n_nuc} nuclei, {s.n_basis} "
                        f"Cartesian s/p/d basis functions, {args.walkers} walkers, "
                        f"single-electron importance-sampled VMC, tau={args.tau}",
            "precision": "FP64 throughout" if args.precision == "fp64" else
                         ({"fp32": "basis FP64, orbital GEMM FP32 (MFMA), derivative matrices FP32",
                           "fp32store": "basis and orbital GEMM FP64, derivative matrices FP32",
                           "bf16": "basis FP64, orbital GEMM BF16 (MFMA, FP32 accumulate), "
                                   "derivative matrices FP32",
                           "bf16store": "basis FP64, orbital GEMM FP32 (MFMA), derivative "
                                        "matrices BF16"}[args.precision]
                          + f"; SM inverse/log-det/E_L FP64, FP64 re-inversion every "
                            f"{args.refresh} moves"),
            "walkers": args.walkers, "electrons": N, "basis_functions": s.n_basis,
            "walker_groups": mixed["walker_groups"],
        },
        "fp64": None if fp64 is None else {
            "value": fp64["walker_moves_per_s"], "ms_per_step": fp64["ms_per_step"],
            "acceptance": fp64["acceptance"], "mean_e_local": fp64["mean_e_local"],
            "mol256_block_latency_ms": block_stats(fp64["block_ms"])},
        "mixed_over_fp64": None if fp64 is None else
        mixed["walker_moves_per_s"] / fp64["walker_moves_per_s"],
        "block_latency_ms": block_stats(mixed["block_ms"]),
        "mol256_block_latency_ms": block_stats(mixed["block_ms"]),
        "acceptance": mixed["acceptance"],
        "mean_e_local": mixed["mean_e_local"],
        "ms_per_step_one_group_with_kernel_events": mixed["ms_per_step_one_group_timed"],
        "roofline": {"kernel": "sm_move", "bound": "hbm", "achieved": achieved,
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                     "traffic": measured_traffic(args.precision)
                     if args.walkers == 8192 and args.config == 2 else None,
                     "traffic_source": "rocprofv3 FETCH_SIZE x2 + WRITE_SIZE x1 per dispatch, "
                                       "profiles/sm_move_mol256.md",
                     "avg_launch_ms": avg_ms,
                     "algorithmic_bytes_per_launch": bytes_launch},
        "kernels_ms_per_step": {k: v["ms"] / max(args.steps, 1) for k, v in mixed["kernels"].items()},
        "init_s": mixed["init_s"],
    }
    if fp64 is not None:
        out["fp64"]["kernels_ms_per_step"] = {k: v["ms"] / max(args.steps, 1)
                                              for k, v in fp64["kernels"].items()}
    if store is not None:
        out["fp32store"] = {
            "description": "FP64 basis + FP64 MFMA GEMM, derivative matrices stored in FP32, "
                           "FP64 inverse/log-det/E_L",
            "value": store["walker_moves_per_s"], "ms_per_step": store["ms_per_step"],
            "over_fp64": store["walker_moves_per_s"] / fp64["walker_moves_per_s"],
            "kernels_ms_per_step": {k: v["ms"] / max(args.steps, 1)
                                    for k, v in store["kernels"].items()}}
    if bf16 is not None:
        out["bf16"] = {
            "description": "FP64 basis rounded to BF16, BF16 coefficients, "
                           "v_mfma_f32_16x16x32_bf16 with FP32 accumulate, FP32 derivative "
                           "storage, FP64 inverse/log-det/E_L",
            "value": bf16["walker_moves_per_s"], "ms_per_step": bf16["ms_per_step"],
            "over_fp64": None if fp64 is None else
            bf16["walker_moves_per_s"] / fp64["walker_moves_per_s"],
            "kernels_ms_per_step": {k: v["ms"] / max(args.steps, 1)
                                    for k, v in bf16["kernels"].items()}}
    if bf16store is not None:
        out["bf16store"] = {
            "description": "FP64 basis rounded to FP32, v_mfma_f32_16x16x4f32 GEMM, derivative "
                           "matrices stored in BF16 (8-bit mantissa: 16 n^2 bytes per move "
                           "against 24 n^2), FP64 inverse/log-det/E_L",
            "value": bf16store["walker_moves_per_s"], "ms_per_step": bf16store["ms_per_step"],
            "over_fp64": None if fp64 is None else
            bf16store["walker_moves_per_s"] / fp64["walker_moves_per_s"],
            "kernels_ms_per_step": {k: v["ms"] / max(args.steps, 1)
                                    for k, v in bf16store["kernels"].items()}}
    if args.accuracy and args.precision != "fp64":
        R_acc, acc_desc = equilibrated_configs(s, args.accuracy)
        out["mixed_error"] = accuracy_report(s, R_acc, acc_desc, precision=args.precision)
        if store is not None:
            out["fp32store"]["error"] = accuracy_report(s, R_acc, acc_desc, precision="fp32store")
        if bf16 is not None:
            out["bf16"]["error"] = accuracy_report(s, R_acc, acc_desc, precision="bf16")
        if bf16store is not None:
            out["bf16store"]["error"] = accuracy_report(s, R_acc, acc_desc, precision="bf16store")
    if args.serving_jobs > 0:
        srv = serving_run(args.precision, args.serving_jobs, args.rate, 20, args.serving_workers,
                          args.capacity_gib)
        srv64 = None if args.no_fp64 else serving_run(
            "fp64", args.serving_jobs, args.rate, 20, args.serving_workers, args.capacity_gib)
        out["serving"] = {
            "workload": f"{args.serving_jobs} VMC/DMC jobs, Poisson arrivals {args.rate}/s, "
                        "8-256 e-, 1e3-1e5 walker-moves per job in blocks of 100 steps, "
                        f"{args.serving_workers} job slots, admission by projected footprint",
            "block_latency_ms": srv["block_latency_ms"],
            "walker_moves_per_s": srv["walker_moves_per_s"], "wall_s": srv["wall_s"],
            "completed": srv["completed"], "failed": srv["failed"], "rejected": srv["rejected"],
            "peak_active": srv["peak_active"], "peak_bytes": srv["peak_bytes"],
            "warmup_jobs": srv["warmup_jobs"],
            "fp64": None if srv64 is None else {
                "block_latency_ms": srv64["block_latency_ms"],
                "walker_moves_per_s": srv64["walker_moves_per_s"], "wall_s": srv64["wall_s"],
                "completed": srv64["completed"], "failed": srv64["failed"]}}
        # the metric's p50/p99 block latency is the serving loop's (blocks of
        # 100 steps of jobs arriving while others run); the mol256 run's
        # blocks are reported beside it
        out["block_latency_ms"] = dict(srv["block_latency_ms"], block_steps=100,
                                       source="serving loop, mixed precision")
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(s, args.tau, 99)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
