"""The FP64 Slater inverse at full batch size: 8192 walkers of the 256-electron
molecule (two 128 x 128 matrices per walker) evaluated at once, so that one
inverse launch holds 8192 matrices.

The inverse enters every evaluate output (ln|Psi| through the determinant,
grad ln Psi, Laplacian(Psi)/Psi and E_L through A^-1 contracted with the
derivative matrices), so walkers spread over the batch are checked against
the C oracle (its LU follows Array2D.h:1204-1421) at the FP64 bar of the
evaluate parity tests: 1e-10 relative (floor 1), sign exact.  A walker with
one electron far from every nucleus has an all-zero row of orbital values
(every basis function underflows): the reference's LU reports that as
singular (calcOK = false, Array2D.h:1221-1235), and so must the batched
inverse, without disturbing its neighbours in the batch.
"""
import sys

import numpy as np
import pytest

from qmcbeaver_amd import synth, Context
from oracle.pyoracle import Oracle

pytestmark = pytest.mark.gpu
W = 8192


def rel(a, b):
    a, b = np.asarray(a, float), np.asarray(b, float)
    return float(np.max(np.abs(a - b) / np.maximum(1.0, np.abs(b))))


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def test_inverse_full_batch_matches_oracle():
    s = synth.make_system(2)
    pk = s.pack()
    assert pk["n_up"] == 128
    R = synth.initial_walkers(s, W, 0x1A5)
    sing_w = 4097
    R[sing_w, 0] = 1.0e4  # every orbital value of electron 0 underflows to 0
    check = [0, 1, 1023, 2048, 4096, 4098, 6143, W - 1]
    say("case inverse full batch: mol256 fp64, walkers", W)
    with Context(pk, precision="fp64", max_walkers=W) as ctx:
        ctx.enable_timing(True)
        res = ctx.evaluate(np.arange(W), R)
        kt = ctx.kernel_times()
        ctx.enable_timing(False)
    inv = kt.get("lu_inverse")
    if inv:
        say("lu_inverse: %.3f ms per launch of %d matrices" % (inv["ms"] / inv["launches"], W))
    assert res["singular"][sing_w]
    assert not np.delete(res["singular"], sing_w).any()
    ref = Oracle(pk).evaluate_all(R[check + [sing_w]])
    assert ref["singular"][-1]
    for k in ("log_psi", "e_local", "kinetic", "potential"):
        e = rel(res[k][check], ref[k][:-1])
        say(k, "max rel err %.2e" % e)
        assert e < 1e-10, (k, e)
    assert rel(res["grad_lnpsi"][check], ref["grad_lnpsi"][:-1]) < 1e-10
    np.testing.assert_array_equal(res["sign_psi"][check], ref["sign_psi"][:-1])
    assert np.isfinite(np.delete(res["log_psi"], sing_w)).all()
