"""Statistical parity of the device VMC with the reference.

The reference's H2 6-31G** example is run by the reference build
(tests/golden/make_golden_vmc.py: the reference evaluate, importance-sampled
single-electron moves, the reference DDDA estimator QMCProperty) and by the
device VMC with the same input, time step, walker count and starting
configurations.  The RNG streams differ, so the two energies must agree
within their combined DDDA error bars: |E_dev - E_ref| < 4 sigma.  The
device run's per-sweep series goes through this library's DDDA estimator,
which is itself pinned to the reference's (tests/test_stats.py).
"""
import json
import os

import numpy as np
import pytest

from qmcbeaver_amd import Context
from qmcbeaver_amd._abi import QMCB_DRIFT_UMRIGAR93_UNEQUAL
from qmcbeaver_amd.stats import Property, series_energy
from tests.inputs_util import GOLDEN, configs, load

pytestmark = pytest.mark.gpu


def test_vmc_energy_matches_reference_within_error_bars():
    with open(os.path.join(GOLDEN, "vmc_h2_631gss.json")) as f:
        g = json.load(f)
    packed, _, _ = load(g["input"])
    W, sweeps, tau = g["walkers"], g["sweeps"], g["tau"]
    R0 = configs(packed, W, g["config_seed"])
    N = R0.shape[1]
    print("case vmc statistics H2 6-31G**, walkers", W, "sweeps", sweeps, flush=True)
    with Context(packed, precision="fp64", max_walkers=W, seed=99) as ctx:
        ctx.evaluate(np.arange(W), R0)
        p = Property()
        for blk in range(4):  # the series of each call is appended in order
            n = sweeps // 4
            ctx.vmc_steps(W, blk * n * N, n * N, tau)
            rows = ctx.sample_series()
            assert rows.shape == (n, 4)
            np.testing.assert_array_equal(rows[:, 3], W)
            np.testing.assert_array_equal(rows[:, 1], W)  # VMC weights are 1
            assert np.all((rows[:, 2] > 0) & (rows[:, 2] <= W))
            series_energy(rows, p)
    s = p.summary()
    ref = g["reference"]
    print("device", s.average, "+-", s.std_dev, "depth", s.decorr_depth,
          "reference", ref["ddda_energy"], "+-", ref["ddda_std"], flush=True)
    assert s.n_samples == sweeps
    assert s.decorr_depth >= 0 and 0 < s.std_dev < 0.01
    comb = np.hypot(s.std_dev, ref["ddda_std"])
    assert abs(s.average - ref["ddda_energy"]) < 4 * comb


def test_series_rows_follow_sweeps_across_calls():
    """Rows are recorded for steps that move electron 0, wherever the call
    boundaries fall."""
    packed, _, _ = load("h2_631gss")
    W = 64
    R0 = configs(packed, W, 5)
    with Context(packed, precision="fp64", max_walkers=W, seed=1) as ctx:
        ctx.evaluate(np.arange(W), R0)
        counts = []
        for first, n in [(0, 1), (1, 4), (5, 3), (8, 0), (8, 7)]:
            ctx.vmc_steps(W, first, n, 0.05)
            counts.append(len(ctx.sample_series()))
        # steps 0 | 1-4 (2, 4) | 5-7 (6) | none | 8-14 (8, 10, 12, 14)
        assert counts == [1, 2, 1, 0, 4]


def test_vmc_modified_drift_matches_reference():
    """The reference's default drift, umrigar93_unequalelectrons, at tau = 0.3
    (reference run: tests/golden/vmc_h2_631gss_qf.json).  At this time step
    the acceptance depends on the drift in the proposal and in the Green's
    function ratio: over these 1.6e7 moves the reference accepts 0.92335
    with the plain drift and 0.92125 with the modified one (binomial error
    about 7e-5).  The device acceptance must match the reference's within
    5e-4, and the energy within 4 combined error bars."""
    with open(os.path.join(GOLDEN, "vmc_h2_631gss_qf.json")) as f:
        g = json.load(f)
    packed, _, _ = load(g["input"])
    W, sweeps, tau = g["walkers"], g["sweeps"], g["tau"]
    R0 = configs(packed, W, g["config_seed"])
    N = R0.shape[1]
    print("case vmc modified drift H2 6-31G**, tau", tau, flush=True)
    with Context(packed, precision="fp64", max_walkers=W, seed=98,
                 drift_mod=QMCB_DRIFT_UMRIGAR93_UNEQUAL, dt=tau) as ctx:
        ctx.evaluate(np.arange(W), R0)
        p = Property()
        acc = 0.0
        for blk in range(4):
            n = sweeps // 4
            st = ctx.vmc_steps(W, blk * n * N, n * N, tau)
            acc += st.sum_accept
            series_energy(ctx.sample_series(), p)
    s = p.summary()
    ref = g["reference"]
    acc /= W * sweeps * N
    print("device", s.average, "+-", s.std_dev, "acceptance", acc, "reference",
          ref["ddda_energy"], "+-", ref["ddda_std"], ref["acceptance"], flush=True)
    assert abs(acc - ref["acceptance"]) < 5e-4
    comb = np.hypot(s.std_dev, ref["ddda_std"])
    assert abs(s.average - ref["ddda_energy"]) < 4 * comb
