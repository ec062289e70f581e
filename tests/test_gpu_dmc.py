"""Device DMC: fixed-node importance-sampled single-electron moves, the
reference's once-per-sweep umrigar93_probability_weighted reweighting,
non-unit-weight branching with on-device walker copies, and E_T feedback.

* After several sweeps with splits and fusions, every live walker's
  resident state still describes its positions (zero move vs fresh evaluate).
* The population is controlled around the target.
* Physics: the synthetic H2 trial function (one doubly occupied orbital) has
  no nodes, so DMC projects onto the exact Born-Oppenheimer ground state at
  R = 1.4 bohr, E0 = -1.174475 Ha (Kolos & Wolniewicz), whatever the quality
  of the random trial function; the VMC energy of that trial function lies
  far above.
"""
import numpy as np
import pytest

from qmcbeaver_amd import synth, Context

pytestmark = pytest.mark.gpu


def rel(a, b):
    a, b = np.asarray(a, float), np.asarray(b, float)
    return float(np.max(np.abs(a - b) / np.maximum(1.0, np.abs(b))))


def test_dmc_branching_keeps_walker_states_consistent():
    s = synth.make_system(1)
    W0, cap = 48, 160
    print("case dmc consistency cfg1", flush=True)
    R0 = synth.initial_walkers(s, W0, 41)
    with Context(s.pack(), precision="fp64", max_walkers=cap, seed=3) as ctx, \
            Context(s.pack(), precision="fp64", max_walkers=cap) as fresh:
        ctx.evaluate(np.arange(W0), R0)
        res0 = fresh.evaluate(np.arange(W0), R0)
        e0 = float(np.mean(res0["e_local"]))
        ctx.dmc_init(W0)
        # aggressive thresholds so that splits and fusions happen in a short run
        st = ctx.dmc_steps(0, 6 * s.n_elec, 0.05, e0, target=W0, branch_threshold=1.05,
                           fusion_threshold=0.95, reweight_after=0)
        print("pop", st.population, "split", st.n_split, "fused", st.n_fused, "killed",
              st.n_killed, "E_T", st.e_trial, flush=True)
        assert st.n_split > 0 and st.n_fused > 0
        slots, w = ctx.dmc_population()
        assert len(slots) == st.population and np.isfinite(w).all() and (w > 0).all()
        R = ctx.positions(cap)[slots]
        ref = fresh.evaluate(slots, R)
        for e in (0, 7, s.n_up + 1):
            got = ctx.evaluate(slots, R, whichE=e)
            assert rel(got["e_local"], ref["e_local"]) < 1e-8, e
            assert rel(got["log_psi"], ref["log_psi"]) < 1e-8
            ctx.accept(slots, np.zeros(len(slots), dtype=np.uint8))


def test_dmc_h2_reaches_exact_ground_state():
    s = synth.make_system(0)
    W = 2048
    print("case dmc H2 exact ground state", flush=True)
    R0 = synth.initial_walkers(s, W, 42)
    with Context(s.pack(), precision="fp64", max_walkers=4 * W, seed=11) as ctx:
        ctx.evaluate(np.arange(W), R0)
        vmc = ctx.vmc_steps(W, 0, 400, 0.05)
        e_vmc = vmc.sum_e_local / vmc.n_samples
        ctx.dmc_init(W)
        tau = 0.02
        st = ctx.dmc_steps(0, 3000, tau, e_vmc, target=W)  # equilibration
        means = []
        for blk in range(8):
            st = ctx.dmc_steps(3000 + 500 * blk, 500, tau, st.e_trial, target=W)
            means.append(st.sum_w_e_local / st.sum_weights)
            assert 0.5 * W < st.population < 2.0 * W
        m = float(np.mean(means))
        err = float(np.std(means) / np.sqrt(len(means)))
        print("VMC", e_vmc, "DMC", m, "+-", err, "blocks", means, flush=True)
        assert abs(m - (-1.174475)) < max(5 * err, 0.01)
        assert e_vmc > m + 0.05
