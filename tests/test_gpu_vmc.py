"""Device VMC (propose -> evaluate at the proposal -> Metropolis -> commit,
with delayed Sherman-Morrison updates and periodic FP64 re-inversion).

* After hundreds of accepted and rejected moves, the resident state (inverse
  with pending delayed updates, log-det, grad/Laplacian traces, Jastrow and
  potential sums) must describe the current positions: a zero-length
  proposal evaluated against it equals a fresh all-electron evaluate of the
  same positions (FP64: 1e-8 relative, limited by SM drift over ~10^3
  updates; mixed: 1e-6).
* Each walker's trajectory depends only on (seed, walker, step): the same
  walkers run alone or in a larger batch end bit-identical.
"""
import os
import sys

import numpy as np
import pytest

from qmcbeaver_amd import synth, Context

pytestmark = pytest.mark.gpu


def rel(a, b):
    a, b = np.asarray(a, float), np.asarray(b, float)
    return float(np.max(np.abs(a - b) / np.maximum(1.0, np.abs(b))))


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("precision,refresh,tol", [("fp64", 0, 1e-8), ("fp32", 64, 1e-6),
                                                    ("bf16", 64, 1e-6), ("bf16store", 64, 1e-6)])
def test_vmc_resident_state_matches_fresh_evaluate(precision, refresh, tol):
    s = synth.make_system(1)
    W, steps = 64, 300
    say("vmc consistency", precision, "walkers", W, "steps", steps)
    R0 = synth.initial_walkers(s, W, 21)
    with Context(s.pack(), precision=precision, max_walkers=W, refresh_interval=refresh,
                 seed=77) as ctx, \
            Context(s.pack(), precision=precision, max_walkers=W) as fresh:
        ids = np.arange(W)
        ctx.evaluate(ids, R0)
        st = ctx.vmc_steps(W, 0, steps, 0.02)
        acc = st.sum_accept / (W * steps)
        say("acceptance", acc, "mean E_L", st.sum_e_local / st.n_samples)
        assert 0.2 < acc < 1.0
        R = ctx.positions(W)
        assert np.isfinite(R).all()
        ref = fresh.evaluate(ids, R)
        assert not ref["singular"].any()
        for e in (0, 5, s.n_up, s.n_elec - 1):
            res = ctx.evaluate(ids, R, whichE=e)  # zero move against the resident state
            assert np.allclose(res["det_ratio"], 1.0, rtol=1e-9, atol=0), e
            assert rel(res["e_local"], ref["e_local"]) < tol, (e, rel(res["e_local"], ref["e_local"]))
            assert rel(res["log_psi"], ref["log_psi"]) < tol
            assert rel(res["grad_lnpsi"], ref["grad_lnpsi"]) < tol
            ctx.accept(ids, np.zeros(W, dtype=np.uint8))


@pytest.mark.parametrize("n_elec", [152, 200], ids=["n76_tile96", "n100_tile128"])
@pytest.mark.parametrize("precision,refresh,tol", [("fp64", 0, 1e-8), ("fp32", 64, 1e-6),
                                                    ("bf16store", 64, 1e-6)])
def test_vmc_resident_state_column_tiles(n_elec, precision, refresh, tol):
    # the same check on the 96- and 128-column Sherman-Morrison tiles
    s = synth.make_system(-1, seed=5, n_elec=n_elec)
    W, steps = 32, 200
    say("vmc consistency", precision, "electrons", n_elec, "walkers", W, "steps", steps)
    R0 = synth.initial_walkers(s, W, 23)
    with Context(s.pack(), precision=precision, max_walkers=W, refresh_interval=refresh,
                 seed=78) as ctx, \
            Context(s.pack(), precision=precision, max_walkers=W) as fresh:
        ids = np.arange(W)
        ctx.evaluate(ids, R0)
        st = ctx.vmc_steps(W, 0, steps, 0.01)
        acc = st.sum_accept / (W * steps)
        say("acceptance", acc)
        assert 0.2 < acc < 1.0
        R = ctx.positions(W)
        ref = fresh.evaluate(ids, R)
        assert not ref["singular"].any()
        for e in (0, s.n_up - 1, s.n_up, s.n_elec - 1):
            res = ctx.evaluate(ids, R, whichE=e)
            assert np.allclose(res["det_ratio"], 1.0, rtol=1e-9, atol=0), e
            assert rel(res["e_local"], ref["e_local"]) < tol, (e, rel(res["e_local"], ref["e_local"]))
            assert rel(res["log_psi"], ref["log_psi"]) < tol
            assert rel(res["grad_lnpsi"], ref["grad_lnpsi"]) < tol
            ctx.accept(ids, np.zeros(W, dtype=np.uint8))


def test_vmc_walkers_independent_of_batch():
    s = synth.make_system(1)
    say("vmc batch independence")
    R0 = synth.initial_walkers(s, 96, 22)
    out = []
    for W in (32, 96):
        with Context(s.pack(), precision="fp32", max_walkers=W, refresh_interval=16,
                     seed=5) as ctx:
            ctx.evaluate(np.arange(W), R0[:W])
            ctx.vmc_steps(W, 0, 80, 0.02)
            out.append(ctx.positions(W)[:32])
    np.testing.assert_array_equal(out[0], out[1])


def test_vmc_deterministic_and_seeded():
    s = synth.make_system(0)
    say("vmc determinism (H2)")
    R0 = synth.initial_walkers(s, 128, 23)
    stats = []
    for seed in (9, 9, 10):
        with Context(s.pack(), precision="fp64", max_walkers=128, seed=seed) as ctx:
            ctx.evaluate(np.arange(128), R0)
            st = ctx.vmc_steps(128, 0, 200, 0.1)
            stats.append((st.sum_e_local, st.sum_accept))
    assert stats[0] == stats[1]
    assert stats[0] != stats[2]


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_vmc_walker_groups_bit_identical(precision):
    """Walker groups on their own streams (overlapping kernels) change the
    schedule, not the arithmetic: positions, energy and acceptance sums and
    the per-sweep series equal the one-group run bit for bit, across sweep
    boundaries and FP64 re-inversions."""
    s = synth.make_system(1)
    W, N = 1024, s.n_elec
    R0 = synth.initial_walkers(s, W, 23)
    out = []
    for groups in (1, 4, 3):
        say("vmc walker groups", precision, groups)
        with Context(s.pack(), precision=precision, max_walkers=W, seed=5,
                     refresh_interval=16 if precision != "fp64" else 0,
                     walker_groups=groups) as ctx:
            ctx.evaluate(np.arange(W), R0)
            st = ctx.vmc_steps(W, 0, N + 9, 0.02)
            rows = ctx.sample_series()
            st2 = ctx.vmc_steps(W, N + 9, N, 0.02)
            rows2 = ctx.sample_series()
            out.append((ctx.positions(W), (st.sum_e_local, st.sum_accept, st.sum_prop_e_local,
                                           st2.sum_e_local, st2.sum_accept), rows, rows2))
    for o in out[1:]:
        np.testing.assert_array_equal(o[0], out[0][0])
        assert o[1] == out[0][1]
        np.testing.assert_array_equal(o[2], out[0][2])
        np.testing.assert_array_equal(o[3], out[0][3])


@pytest.mark.parametrize("W,refresh,steps", [(1024, 64, 200), (8192, 256, 320)])
def test_vmc_walker_groups_chosen_by_timing(W, refresh, steps):
    """walker_groups = -1: the first long call times windows of 1, 2 and 3
    groups inside its own steps (an untimed round, then two timed rounds)
    and keeps the fastest; the trajectory, the call's sums and the
    per-sweep series equal the one-group run bit for bit, and the next call
    keeps the choice.  At 8192 walkers and a re-inversion every 256 moves the
    windows are one refresh phase (32 steps) long."""
    s = synth.make_system(1)
    N = s.n_elec
    R0 = synth.initial_walkers(s, W, 29)
    out = []
    for groups in (1, -1):
        say("vmc walker groups", groups)
        with Context(s.pack(), precision="fp32", max_walkers=W, seed=6, refresh_interval=refresh,
                     walker_groups=groups) as ctx:
            ctx.evaluate(np.arange(W), R0)
            assert ctx.walker_groups() == 0  # no steps yet
            st = ctx.vmc_steps(W, 0, steps, 0.02)
            chosen = ctx.walker_groups()
            st2 = ctx.vmc_steps(W, steps, N, 0.02)
            out.append((ctx.positions(W), (st.sum_e_local, st.sum_accept, st.sum_prop_e_local,
                                           st2.sum_e_local, st2.sum_accept),
                        ctx.sample_series(), chosen, ctx.walker_groups()))
    say("chosen walker groups", out[1][3])
    assert out[1][3] in (1, 2, 3) and out[1][4] == out[1][3]
    np.testing.assert_array_equal(out[1][0], out[0][0])
    assert out[1][1] == out[0][1]
    np.testing.assert_array_equal(out[1][2], out[0][2])


@pytest.mark.parametrize("W,K,S", [(4096, 16, 4), (256, 16, 16)])
def test_vmc_refresh_is_staggered(W, K, S):
    """refresh_interval K: the slot residues s mod K are grouped into phases
    of S residues, S the shortest phase whose refresh launch carries at least
    1024 walkers (S = ceil(16 * 1024 / W) capped at K); phase q is refreshed
    (FP64 re-inversion and traces) after the move c of the spin with c mod K
    = min(qS + S - 1, K - 1).  20 steps of spin-up moves (n_up = 32) then
    refresh at c = 3, 7, 11, 15, 19 for W = 4096 and at c = 15 for W = 256,
    one re-inversion and one trace launch each."""
    s = synth.make_system(1)
    steps = 20
    say("vmc staggered refresh", W, K)
    expect = sum(1 for c in range(steps) if c % K == min((c % K) // S * S + S - 1, K - 1))
    R0 = synth.initial_walkers(s, W, 24)
    with Context(s.pack(), precision="fp32", max_walkers=W, refresh_interval=K,
                 seed=8) as ctx:
        ctx.evaluate(np.arange(W), R0)
        ctx.enable_timing(True)
        ctx.vmc_steps(W, 0, steps, 0.02)
        kt = ctx.kernel_times()
        ctx.enable_timing(False)
    say("kernels", {k: v["launches"] for k, v in kt.items()})
    assert expect == {4: 5, 16: 1}[S]
    assert kt["lu_inverse"]["launches"] == expect
    assert kt["traces_all"]["launches"] == expect
    assert kt["sm_move"]["launches"] == steps


@pytest.mark.parametrize("config,precision,W,groups", [(1, "fp64", 256, 0), (1, "fp32", 1024, 2),
                                                       (2, "fp32", 512, 0)])
def test_vmc_next_proposal_in_sm_move_bit_identical(config, precision, W, groups, monkeypatch):
    """A step's sm_move draws the next step's proposal (one propose launch per
    call, walkers that a re-inversion touched proposed again); the walk equals
    the one with a propose launch per step bit for bit, with the staggered
    FP64 re-inversion and walker groups."""
    s = synth.make_system(config)
    R0 = synth.initial_walkers(s, W, 31)
    print("case fused proposal config", config, precision, W, "groups", groups, flush=True)
    out = []
    for separate in ("0", "1"):
        monkeypatch.setenv("QMCB_SEPARATE_PROPOSE", separate)
        with Context(s.pack(), precision=precision, max_walkers=W, seed=17,
                     refresh_interval=64 if precision != "fp64" else 0,
                     walker_groups=groups) as c:
            c.evaluate(np.arange(W), R0)
            rec = []
            for call in range(2):
                st = c.vmc_steps(W, 150 * call, 150, 0.02)
                rec.append((c.positions(W), c.sample_series(), st.sum_accept, st.sum_e_local))
            out.append(rec)
    for a, b in zip(*out):
        np.testing.assert_array_equal(a[0], b[0])
        np.testing.assert_array_equal(a[1], b[1])
        assert a[2] == b[2] and a[3] == b[3]


def unequal_spins(n_up=68, n_dn=12, seed=4):
    """The 300-function organic system with 68 + 12 electrons: moving an
    up-spin electron (96-column tile, 192 threads) forms the Jastrow terms
    inside sm_move, moving a down-spin one (16-column tile, 64 threads < 80
    electrons) takes the separate jastrow_move launch."""
    p = dict(synth.make_system(1).pack())
    rng = np.random.default_rng(seed)
    q, _ = np.linalg.qr(rng.normal(size=(p["n_basis"], n_up + n_dn)))
    p["n_up"], p["n_dn"] = n_up, n_dn
    p["mo_up"] = np.ascontiguousarray(q[:, :n_up].T)
    p["mo_dn"] = np.ascontiguousarray(q[:, n_up:].T)
    return p


def test_unequal_spins_evaluate_and_walk():
    """68 + 12 electrons: the FP64 all-electron evaluate against the oracle,
    and a mixed-precision VMC (re-inversions, both Jastrow paths, the next
    proposal drawn in the move kernel) whose resident state matches a fresh
    evaluate and whose walk equals the one with a propose launch per step."""
    from oracle.pyoracle import Oracle
    p = unequal_spins()
    N, W = 80, 256
    rng = np.random.default_rng(8)
    R0 = p["nuc_xyz"][rng.integers(0, p["n_nuc"], size=(W, N))] + 0.8 * rng.normal(size=(W, N, 3))
    say("case unequal spins 68+12")
    with Context(p, precision="fp64", max_walkers=W) as ctx:
        a = ctx.evaluate(np.arange(W), R0)
    o = Oracle(p).evaluate_all(R0[:8])
    for k in ("log_psi", "e_local", "grad_lnpsi", "lap_ratio"):
        err = np.max(np.abs(a.a[k][:8] - o.a[k]) / np.maximum(1.0, np.abs(o.a[k])))
        assert err < 1e-10, (k, err)
    out = []
    for separate in ("0", "1"):
        os.environ["QMCB_SEPARATE_PROPOSE"] = separate
        try:
            with Context(p, precision="fp32", max_walkers=W, seed=3, refresh_interval=16) as ctx:
                ctx.evaluate(np.arange(W), R0)
                st = ctx.vmc_steps(W, 0, 3 * N + 7, 0.02)
                R = ctx.positions(W)
                res = ctx.evaluate(np.arange(W), R, whichE=N - 1)  # zero-length proposal
                out.append((R, ctx.sample_series(), st.sum_accept, res.a["e_local"].copy(),
                            res.a["log_psi"].copy()))
        finally:
            del os.environ["QMCB_SEPARATE_PROPOSE"]
    for x, y in zip(out[0][:3], out[1][:3]):
        np.testing.assert_array_equal(x, y)
    assert 0.05 < out[0][2] / (W * (3 * N + 7)) < 0.99
    with Context(p, precision="fp32", max_walkers=W) as fresh:
        ref = fresh.evaluate(np.arange(W), out[0][0])
    for k, i in (("e_local", 3), ("log_psi", 4)):
        err = np.max(np.abs(out[0][i] - ref.a[k]) / np.maximum(1.0, np.abs(ref.a[k])))
        assert err < 1e-3, (k, err)


@pytest.mark.parametrize("refresh", [0, 256])
def test_mixed_resident_state_drift_256_electrons(refresh):
    """The 256-electron molecule, FP32 derivative storage: after 12 sweeps of
    single-electron moves with no FP64 re-inversion (the reference's own
    single-electron mode) or one every 256 moves of a spin, the resident
    state (Sherman-Morrison-updated FP64 inverse and delayed terms) gives
    the local energy, ln|Psi| and grad ln Psi of a fresh evaluate of the same
    positions to far better than the 1e-6 Ha per electron target."""
    s = synth.make_system(2)
    W, N = 512, s.n_elec
    say("case resident drift, 256 electrons, refresh", refresh)
    R0 = synth.initial_walkers(s, W, 5)
    ids = np.arange(W)
    with Context(s.pack(), precision="fp32", max_walkers=W, refresh_interval=refresh,
                 seed=9) as ctx, \
            Context(s.pack(), precision="fp32", max_walkers=W) as fresh:
        ctx.evaluate(ids, R0)
        done = 0
        for chunk in (256, 2816):
            ctx.vmc_steps(W, done, chunk, 0.01)
            done += chunk
            R = ctx.positions(W)
            ref = fresh.evaluate(ids, R)
            assert not ref["singular"].any()
            res = ctx.evaluate(ids, R, whichE=N - 1)  # zero-length move of the last electron
            errs = {k: rel(res[k], ref[k]) for k in ("e_local", "log_psi", "grad_lnpsi")}
            say("steps", done, errs)
            assert errs["e_local"] < 1e-9 and errs["log_psi"] < 1e-10, errs
            assert errs["grad_lnpsi"] < 1e-7, errs
            ctx.accept(ids, np.zeros(W, dtype=np.uint8))
