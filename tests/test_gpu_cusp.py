"""Cusp replacement on the device against the reference build's golden
vectors (tests/golden/make_golden_cusp.py): configurations with electrons
inside the correction radii, all-electron and along single-electron move
sequences that step into and out of them.  FP64 bar: 1e-10 relative
(all-electron), 1e-9 (Sherman-Morrison sequences)."""
import os
import sys

import numpy as np
import pytest

from qmcbeaver_amd import synth, Context
from tests.inputs_util import GOLDEN, load

pytestmark = pytest.mark.gpu
CASES = ["h2_631gss", "hmx", "cfg1"]


def rel(a, b):
    a, b = np.asarray(a, float), np.asarray(b, float)
    return float(np.max(np.abs(a - b) / np.maximum(1.0, np.abs(b))))


def packed_of(name):
    if name == "cfg1":
        return synth.make_system(1).pack()
    return load(name)[0]


@pytest.mark.parametrize("name", CASES)
def test_cusp_all_electron_matches_reference(name):
    g = np.load(os.path.join(GOLDEN, f"cusp_{name}.npz"))
    pk = packed_of(name)
    R = g["R"]
    print("case cusp all-electron", name, "walkers", R.shape[0], flush=True)
    with Context(pk, precision="fp64", max_walkers=R.shape[0], replace_cusps=True) as ctx:
        res = ctx.evaluate(np.arange(R.shape[0]), R)
    assert not res["singular"].any()
    for k in ("log_psi", "e_local", "kinetic", "v_ne", "v_ee"):
        assert rel(res[k], g["ev_" + k]) < 1e-10, (k, rel(res[k], g["ev_" + k]))
    assert rel(res["grad_lnpsi"], g["ev_grad_lnpsi"]) < 1e-10
    np.testing.assert_array_equal(res["sign_psi"], g["ev_sign_psi"].astype(np.int8))


@pytest.mark.parametrize("name", ["h2_631gss", "cfg1"])
def test_cusp_single_electron_sequence_matches_reference(name):
    """(No HMX sequence: the reference's single-electron row copy reads with
    the wrong stride when, as in that input, there are more orbitals than
    electrons.)"""
    g = np.load(os.path.join(GOLDEN, f"cusp_{name}.npz"))
    pk = packed_of(name)
    mv = g["moves"]
    nw, nm = mv.shape[:2]
    R = g["R"][:nw].copy()
    nrec = nm + 1
    print("case cusp single-electron", name, nw, "walkers", nm, "moves", flush=True)
    with Context(pk, precision="fp64", max_walkers=nw, replace_cusps=True) as ctx:
        ids = np.arange(nw)
        res = ctx.evaluate(ids, R)
        assert rel(res["e_local"], g["sq_e_local"][np.arange(nw) * nrec]) < 1e-10
        for m in range(nm):
            for w in range(nw):
                e = int(mv[w, m, 0])
                assert e == int(mv[0, m, 0])
            e = int(mv[0, m, 0])
            R[:, e, :] = mv[:, m, 1:]
            res = ctx.evaluate(ids, R, whichE=e)
            idx = np.arange(nw) * nrec + m + 1
            assert rel(res["log_psi"], g["sq_log_psi"][idx]) < 1e-9, m
            assert rel(res["e_local"], g["sq_e_local"][idx]) < 1e-9, m
            assert rel(res["grad_lnpsi"], g["sq_grad_lnpsi"][idx]) < 1e-9, m
            ctx.accept(ids, np.ones(nw, dtype=np.uint8))


def test_cusp_mixed_precision_envelope():
    """FP32 derivative storage near nuclei: these configurations put
    electrons inside the correction radii, where |E_L| reaches 10^3-10^4 Ha,
    so the bar is relative: |dE_L| <= 1e-6 max(|E_L|, N)."""
    g = np.load(os.path.join(GOLDEN, "cusp_cfg1.npz"))
    pk = packed_of("cfg1")
    R = g["R"]
    N = R.shape[1]
    print("case cusp fp32store cfg1", flush=True)
    with Context(pk, precision="fp32store", max_walkers=R.shape[0], replace_cusps=True) as ctx:
        res = ctx.evaluate(np.arange(R.shape[0]), R)
    el = np.abs(res["e_local"] - g["ev_e_local"]) / np.maximum(np.abs(g["ev_e_local"]), N)
    assert el.max() < 1e-6, el.max()
