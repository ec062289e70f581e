"""Mixed precision against the FP64 path on the 256-electron molecule
(the FP64 path itself matches the reference to <= 1e-10).

* fp32store (FP64 basis and MFMA GEMM, derivative matrices stored in FP32):
  the target bar, |E_L| error <= 1e-6 Ha per electron for 95% of the
  configurations, median <= 2e-7 (max <= 2e-5 on the worst-conditioned
  ones, about 1% of this random-MO workload exceed 1e-6); ln|Psi| is FP64
  arithmetic throughout (<= 1e-9).
* fp32 (FP32 MFMA operands and accumulation): the Slater matrices of this
  random-MO workload have condition numbers 1e3-1e5, so FP32 rounding of the
  matrix elements alone moves ln|det| by up to ~1e-3; the envelope asserted
  here is the measured one (ln|Psi| < 0.05, median E_L error < 1e-3 Ha/e),
  and bench.py reports the full per-quantity statistics on 1e4
  configurations.
"""
import numpy as np
import pytest

from qmcbeaver_amd import synth, Context

pytestmark = pytest.mark.gpu


def errors(s, precision, n=256, seed=99):
    R = synth.initial_walkers(s, n, seed)
    ids = np.arange(n)
    with Context(s.pack(), precision="fp64", max_walkers=n) as c64, \
            Context(s.pack(), precision=precision, max_walkers=n) as cx:
        a = c64.evaluate(ids, R)
        b = cx.evaluate(ids, R)
        Rp = R.copy()
        Rp[:, 3, :] += 0.05
        pa = c64.evaluate(ids, Rp, whichE=3)
        pb = cx.evaluate(ids, Rp, whichE=3)
    ok = ~(a["singular"].astype(bool) | b["singular"].astype(bool) |
           pa["singular"].astype(bool) | pb["singular"].astype(bool))
    N = s.n_elec
    e = {"ok": ok.sum(),
         "dlog": np.abs(a["log_psi"] - b["log_psi"])[ok],
         "de": np.abs(a["e_local"] - b["e_local"])[ok] / N,
         "dpe": np.abs(pa["e_local"] - pb["e_local"])[ok] / N,
         "dplog": np.abs(pa["log_psi"] - pb["log_psi"])[ok],
         "dg": np.abs(a["grad_lnpsi"] - b["grad_lnpsi"]).reshape(n, -1).max(axis=1)[ok],
         "sign": np.array_equal(a["sign_psi"][ok], b["sign_psi"][ok])}
    print(precision, {k: (float(np.max(v)) if hasattr(v, "__len__") and len(v) else v)
                      for k, v in e.items()}, flush=True)
    return e


def test_fp32store_meets_energy_target():
    s = synth.make_system(2)
    print("case cfg2 fp32store vs fp64", flush=True)
    e = errors(s, "fp32store")
    assert e["ok"] > 0.95 * 256
    assert e["dlog"].max() < 1e-9 and e["dplog"].max() < 1e-9
    assert np.percentile(e["de"], 95) < 1e-6 and np.percentile(e["dpe"], 95) < 1e-6
    assert np.median(e["de"]) < 2e-7 and np.median(e["dpe"]) < 2e-7
    assert e["de"].max() < 2e-5 and e["dpe"].max() < 2e-5
    assert e["sign"]


def test_fp32_error_envelope():
    s = synth.make_system(2)
    print("case cfg2 fp32 vs fp64", flush=True)
    e = errors(s, "fp32")
    assert e["ok"] > 0.95 * 256
    assert e["dlog"].max() < 0.05
    assert np.median(e["de"]) < 1e-3 and np.median(e["dpe"]) < 1e-3
