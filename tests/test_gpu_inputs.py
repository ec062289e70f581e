"""HIP path on the reference's own inputs (H2 6-31G** example, 152-electron
HMX) against the reference build's golden vectors: <= 1e-10 relative."""
import numpy as np
import pytest

from qmcbeaver_amd import Context
from tests.inputs_util import INPUTS, load

pytestmark = pytest.mark.gpu


def rel(a, b):
    a, b = np.asarray(a, float), np.asarray(b, float)
    return float(np.max(np.abs(a - b) / np.maximum(1.0, np.abs(b))))


@pytest.mark.parametrize("name", INPUTS)
def test_reference_input_fp64_matches_reference(name):
    packed, g, flags = load(name)
    R = g["R"]
    print("case input", name, "walkers", R.shape[0], flush=True)
    with Context(packed, precision="fp64", max_walkers=R.shape[0]) as ctx:
        res = ctx.evaluate(np.arange(R.shape[0]), R)
    for k in ("log_psi", "e_local", "kinetic", "v_ne", "v_ee"):
        assert rel(res[k], g["ev_" + k]) < 1e-10, k
    assert rel(res["grad_lnpsi"], g["ev_grad_lnpsi"]) < 1e-10
    assert np.array_equal(res["sign_psi"], g["ev_sign_psi"].astype(np.int8))
    with Context(packed, precision="fp64", max_walkers=R.shape[0], drift_mod=1,
                 dt=float(g["dt"])) as ctx:
        res = ctx.evaluate(np.arange(R.shape[0]), R)
    assert rel(res["mod_drift"], g["umr_mod_drift"]) < 1e-10
