"""The .ckmf reader maps the reference's own inputs (H2 6-31G** example, the
152-electron HMX regression input) onto the same system the reference reads:
the oracle evaluating the reader's output matches the reference build's
golden vectors on the same file."""
import numpy as np
import pytest

from oracle.pyoracle import Oracle
from tests.inputs_util import INPUTS, load


def rel(a, b):
    a, b = np.asarray(a, float), np.asarray(b, float)
    return float(np.max(np.abs(a - b) / np.maximum(1.0, np.abs(b))))


@pytest.mark.parametrize("name", INPUTS)
def test_reader_plus_oracle_match_reference(name):
    packed, g, flags = load(name)
    R = g["R"]
    n = min(R.shape[0], 4)
    o = Oracle(packed)
    res = o.evaluate_all(R[:n])
    for k in ("log_psi", "e_local", "kinetic", "v_ne", "v_ee"):
        assert rel(res[k], g["ev_" + k][:n]) < 1e-10, k
    assert rel(res["grad_lnpsi"], g["ev_grad_lnpsi"][:n]) < 1e-10
    o2 = Oracle(packed, drift_mod=1, dt=float(g["dt"]))
    res2 = o2.evaluate_all(R[:n])
    assert rel(res2["mod_drift"], g["umr_mod_drift"][:n]) < 1e-10


def test_reader_shapes():
    packed, g, flags = load("hmx")
    assert packed["n_up"] == 76 and packed["n_dn"] == 76
    assert packed["n_basis"] == 516 and packed["n_nuc"] == 28
    assert packed["mo_up"].shape == (76, 516)


def test_reader_multideterminant_tables():
    """The Be MCSCF example: 10 determinants over 5 orbitals; each
    determinant's columns are its occupied orbitals in orbital order among the
    active ones (QMCWavefunction::makeCoefficients)."""
    import gzip, os, tempfile
    from qmcbeaver_amd import ckmf
    from tests.inputs_util import GOLDEN
    with tempfile.TemporaryDirectory() as tmp:
        raw = os.path.join(tmp, "be.ckmf")
        with gzip.open(os.path.join(GOLDEN, "be_mcscf.ckmf.gz"), "rb") as g, open(raw, "wb") as f:
            f.write(g.read())
        pk, flags, info = ckmf.read_ckmf(raw)
    assert pk["n_det"] == 10 and pk["n_up"] == 2 and pk["n_dn"] == 2
    assert pk["n_act_up"] == 5 and pk["mo_up"].shape == (5, 9)
    assert pk["occ_up"].tolist()[:4] == [[0, 1], [0, 3], [0, 2], [0, 4]]
    assert pk["occ_dn"].tolist()[4] == [0, 4]
    assert abs(pk["ci_coef"][0] - 0.9497089) < 1e-12
    assert abs(pk["ci_coef"].sum() - (0.9497089 - 3 * 0.1807882 + 0.0000002)) < 1e-6
