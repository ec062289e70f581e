"""Orbital counts that are not multiples of the 16-orbital MFMA tile.

Within a full 16x16 tile any mix-up of the accumulator-to-orbital mapping is
an even permutation of the orbitals, which leaves |det|, its sign and every
trace unchanged; a partial last tile turns such a mix-up into a singular
Slater matrix.  These cases pin the mapping for both MFMA instructions
(FP64 and FP32) against the oracle: all-electron evaluate and a sequence of
single-electron proposals (the row GEMM and its LDS-transposed epilogue).

Bars: FP64 1e-10 relative (1e-9 after Sherman-Morrison updates); mixed
modes within their stated envelopes (FP32 store: E_L 1e-6 Ha/electron;
FP32 operands: ln|Psi| 0.05, median E_L 1e-3 Ha/electron -- single
ill-conditioned configurations can be far worse in that mode; BF16 storage
of the derivatives: ln|Psi| as FP32 operands, median E_L 5e-2 on these
unequilibrated configurations).  BF16
operands (the third MFMA instruction, 16x16x16) carry 8-bit mantissas, so
on these unequilibrated configurations ln|Psi| moves by up to ~2 and E_L by
up to tens of Ha/electron; the pin there is the mapping itself: no
singular matrix, finite results, ln|Psi| within 5.
"""
import sys

import numpy as np
import pytest

from qmcbeaver_amd import synth, Context
from oracle.pyoracle import Oracle

pytestmark = pytest.mark.gpu
SIZES = [40, 80, 100, 152]  # orbitals per spin: 20, 40, 50, 76


def rel(a, b):
    a, b = np.asarray(a, float), np.asarray(b, float)
    return float(np.max(np.abs(a - b) / np.maximum(1.0, np.abs(b))))


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module", params=SIZES, ids=[f"mol{n}" for n in SIZES])
def system(request):
    s = synth.make_system(-1, n_elec=request.param)
    return s, s.pack()


def test_all_electron_fp64_partial_tiles(system):
    s, pk = system
    R = synth.initial_walkers(s, 3, 31)
    say("case partial tile fp64 all-electron, n_up", pk["n_up"])
    ref = Oracle(pk).evaluate_all(R)
    with Context(pk, precision="fp64", max_walkers=3) as ctx:
        res = ctx.evaluate(np.arange(3), R)
    assert not res["singular"].any() and not ref["singular"].any()
    for k in ("log_psi", "e_local", "kinetic", "potential"):
        assert rel(res[k], ref[k]) < 1e-10, (k, rel(res[k], ref[k]))
    assert rel(res["grad_lnpsi"], ref["grad_lnpsi"]) < 1e-10
    np.testing.assert_array_equal(res["sign_psi"], ref["sign_psi"])


def test_single_electron_partial_tiles(system):
    s, pk = system
    W = 2
    R = synth.initial_walkers(s, W, 32)
    o = Oracle(pk)
    ws = [o.new_walker() for _ in range(W)]
    for w in range(W):
        o.evaluate(ws[w], R[w], -1)
    say("case partial tile fp64 single-electron, n_up", pk["n_up"])
    rng = np.random.default_rng(8)
    N = s.n_elec
    with Context(pk, precision="fp64", max_walkers=W) as ctx:
        ids = np.arange(W)
        ctx.evaluate(ids, R)
        for m in range(12):
            e = (m * 13) % N
            R[:, e, :] += 0.2 * rng.normal(size=(W, 3))
            res = ctx.evaluate(ids, R, whichE=e)
            for w in range(W):
                ref = o.evaluate(ws[w], R[w], e)
                assert rel(res["log_psi"][w], ref["log_psi"][0]) < 1e-9, (m, e)
                assert rel(res["e_local"][w], ref["e_local"][0]) < 1e-9, (m, e)
                assert rel(res["grad_lnpsi"][w], ref["grad_lnpsi"][0]) < 1e-9, (m, e)
            ctx.accept(ids, np.ones(W, dtype=np.uint8))


@pytest.mark.parametrize("precision", ["fp32store", "fp32", "bf16", "bf16store"])
def test_mixed_partial_tiles(system, precision):
    s, pk = system
    W = 3
    R = synth.initial_walkers(s, W, 33)
    N = s.n_elec
    say("case partial tile", precision, "n_up", pk["n_up"])
    ref = Oracle(pk).evaluate_all(R)
    with Context(pk, precision=precision, max_walkers=W) as ctx:
        res = ctx.evaluate(np.arange(W), R)
        assert not res["singular"].any()
        el = np.abs(res["e_local"] - ref["e_local"]) / N
        dl = np.abs(res["log_psi"] - ref["log_psi"])
        if precision == "fp32store":
            assert el.max() < 1e-6 and dl.max() < 1e-9, (el.max(), dl.max())
        elif precision == "fp32":
            assert np.median(el) < 1e-3 and dl.max() < 0.05, (el, dl.max())
        elif precision == "bf16store":
            assert np.median(el) < 5e-2 and dl.max() < 0.05, (el, dl.max())
        else:
            assert np.isfinite(res["e_local"]).all() and dl.max() < 5.0, (el, dl)
        # one single-electron move through the FP32/FP64 row GEMM epilogue
        e = N - 1
        R2 = R.copy()
        R2[:, e, :] += 0.1
        res2 = ctx.evaluate(np.arange(W), R2, whichE=e)
    ref2 = Oracle(pk).evaluate_all(R2)
    el2 = np.abs(res2["e_local"] - ref2["e_local"]) / N
    if precision == "fp32store":
        assert el2.max() < 1e-6, el2.max()
    elif precision == "fp32":
        assert np.median(el2) < 1e-3, el2
    elif precision == "bf16store":
        assert np.median(el2) < 5e-2, el2
    else:
        assert not res2["singular"].any() and np.isfinite(res2["e_local"]).all()
        assert np.abs(res2["log_psi"] - ref2["log_psi"]).max() < 5.0
