"""Mixed precision on the 256-electron molecule, pinned at the accuracy target.

Configurations: an equilibrated FP64 device VMC with the reference's default
proposal (Umrigar-93 modified drift for unequal electrons, QMCFlags.cpp:
77-78), tau = 0.3, 1024 walkers after 80 sweeps (the mean local energy and
the acceptance are flat from about sweep 60; no electron strays beyond
~15 bohr of the nuclei).  Each mode is compared with the FP64 path (which
matches the reference evaluate to <= 1e-10) on all-electron evaluates and on
a single-electron proposal, and on the four reference golden configurations
directly against the values the reference program wrote.

* fp32store (FP64 basis and GEMM, derivative matrices stored in FP32): the
  target, |E_L| error <= 1e-6 Ha per electron on every configuration and
  on every golden configuration; ln|Psi| is FP64 arithmetic (<= 1e-9).
* fp32 (FP32 MFMA operands and accumulation): measured envelope, median
  E_L error ~3e-7 Ha/e, p99 ~3e-6, so it misses the target on ~10% of
  configurations; asserted with a factor of ~5 margin.
* bf16 (BF16 MFMA operands, FP32 accumulation): median ~1e-3 Ha/e, p99
  ~1e-2 Ha/e; a throughput mode, asserted at its measured envelope.
* bf16store (FP32 MFMA GEMM, derivative matrices stored in BF16): median
  ~1e-4 Ha/e, p99 ~4e-4; the mode that streams 16 n^2 bytes per move
  instead of 24 n^2, asserted at its measured envelope.
"""
import os

import numpy as np
import pytest

from qmcbeaver_amd import synth, Context
from qmcbeaver_amd._abi import QMCB_DRIFT_UMRIGAR93_UNEQUAL

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(__file__), "golden")
W, SWEEPS, TAU = 1024, 80, 0.3


@pytest.fixture(scope="module")
def equilibrated():
    s = synth.make_system(2)
    pk = s.pack()
    N = s.n_elec
    print("case mixed precision: equilibrating %d walkers, %d sweeps, tau %g" % (W, SWEEPS, TAU),
          flush=True)
    ids = np.arange(W)
    with Context(pk, precision="fp64", max_walkers=W, seed=0x5EED,
                 drift_mod=QMCB_DRIFT_UMRIGAR93_UNEQUAL, dt=TAU) as c:
        c.evaluate(ids, synth.initial_walkers(s, W, 0x5EED))
        c.vmc_steps(W, 0, SWEEPS * N, TAU)
        R = c.positions(W)
    Rp = R.copy()
    Rp[:, 5, :] += 0.05
    with Context(pk, precision="fp64", max_walkers=W) as c:
        a = c.evaluate(ids, R)
        p = c.evaluate(ids, Rp, whichE=5)
    return s, pk, R, Rp, a, p


def errors(eq, precision):
    s, pk, R, Rp, a, p = eq
    N = s.n_elec
    ids = np.arange(W)
    with Context(pk, precision=precision, max_walkers=W) as c:
        b = c.evaluate(ids, R)
        pb = c.evaluate(ids, Rp, whichE=5)
    ok = ~(a["singular"].astype(bool) | b["singular"].astype(bool) |
           p["singular"].astype(bool) | pb["singular"].astype(bool))
    e = {"ok": int(ok.sum()),
         "de": np.abs(a["e_local"] - b["e_local"])[ok] / N,
         "dpe": np.abs(p["e_local"] - pb["e_local"])[ok] / N,
         "dlog": np.abs(a["log_psi"] - b["log_psi"])[ok],
         "dplog": np.abs(p["log_psi"] - pb["log_psi"])[ok],
         "sign_mismatch": int(np.sum(a["sign_psi"][ok] != b["sign_psi"][ok]))}
    print(precision, {k: ("med %.2e p99 %.2e max %.2e" % (np.median(v), np.percentile(v, 99), v.max())
                          if isinstance(v, np.ndarray) else v) for k, v in e.items()}, flush=True)
    return e


def golden_errors(precision):
    g = np.load(os.path.join(GOLDEN, "cfg2_mol256.npz"))
    s = synth.make_system(2)
    assert s.digest() == str(g["digest"])
    R = g["R"]
    with Context(s.pack(), precision=precision, max_walkers=len(R)) as c:
        r = c.evaluate(np.arange(len(R)), R)
    assert not r["singular"].any()
    de = np.abs(r["e_local"] - g["ev_e_local"]) / s.n_elec
    dlog = np.abs(r["log_psi"] - g["ev_log_psi"]) / np.maximum(1.0, np.abs(g["ev_log_psi"]))
    print(precision, "vs reference goldens: E_L/e", de, "ln|Psi| rel", dlog, flush=True)
    return de, dlog, np.array_equal(r["sign_psi"], g["ev_sign_psi"])


def test_fp32store_meets_energy_target(equilibrated):
    e = errors(equilibrated, "fp32store")
    assert e["ok"] == W
    assert e["de"].max() <= 1e-6 and e["dpe"].max() <= 1e-6
    assert e["dlog"].max() < 1e-9 and e["dplog"].max() < 1e-9
    assert e["sign_mismatch"] == 0


def test_fp32store_against_reference_goldens():
    de, dlog, sign_ok = golden_errors("fp32store")
    assert de.max() <= 1e-6
    assert dlog.max() < 1e-10
    assert sign_ok


def test_fp32_measured_envelope(equilibrated):
    e = errors(equilibrated, "fp32")
    assert e["ok"] >= W - 8
    assert np.median(e["de"]) < 2e-6 and np.median(e["dpe"]) < 2e-6
    assert np.percentile(e["de"], 99) < 2e-5 and np.percentile(e["dpe"], 99) < 2e-5
    assert e["de"].max() < 2e-4 and e["dpe"].max() < 2e-4
    assert e["dlog"].max() < 2e-3
    de, dlog, _ = golden_errors("fp32")
    assert de.max() < 2e-3 and dlog.max() < 1e-5


def test_bf16_measured_envelope(equilibrated):
    e = errors(equilibrated, "bf16")
    assert e["ok"] >= W - 8
    assert np.median(e["de"]) < 5e-3 and np.percentile(e["de"], 99) < 5e-2
    assert np.median(e["dpe"]) < 5e-3 and np.percentile(e["dpe"], 99) < 5e-2
    assert np.median(e["dlog"]) < 0.25
    assert e["sign_mismatch"] <= W // 100
    de, dlog, _ = golden_errors("bf16")
    assert np.isfinite(de).all() and de.max() < 10.0


def test_bf16store_measured_envelope(equilibrated):
    e = errors(equilibrated, "bf16store")
    assert e["ok"] >= W - 8
    assert np.median(e["de"]) < 3e-4 and np.percentile(e["de"], 99) < 2e-3
    assert np.median(e["dpe"]) < 3e-4 and np.percentile(e["dpe"], 99) < 2e-3
    assert e["de"].max() < 3e-2 and e["dpe"].max() < 3e-2
    assert e["dlog"].max() < 2e-3
    assert e["sign_mismatch"] == 0
    de, dlog, sign_ok = golden_errors("bf16store")
    assert de.max() < 1e-2 and dlog.max() < 1e-5 and sign_ok
