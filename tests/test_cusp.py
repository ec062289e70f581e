"""Cusp-replacement fit (host code of libqmcb_amd.so, no device) against the
parameters the reference build prints for the same inputs
(tests/golden/make_golden_cusp.py): the H2 6-31G** example, the 152-electron
HMX input and the synthetic 64-electron molecule.  Radii, shift, sign, the
quartic's coefficients, phi0, n0, Zeff, the deviation it minimised and the
ideal curve must agree to 1e-9 relative; replacement radii exactly where the
fit is skipped."""
import os

import numpy as np
import pytest

from qmcbeaver_amd import synth
from qmcbeaver_amd.core import cusp_fit
from tests.inputs_util import GOLDEN, load

CASES = ["h2_631gss", "hmx", "cfg1"]


def packed_of(name):
    if name == "cfg1":
        return synth.make_system(1).pack()
    return load(name)[0]


@pytest.mark.parametrize("name", CASES)
def test_cusp_fit_matches_reference(name):
    g = np.load(os.path.join(GOLDEN, f"cusp_{name}.npz"))
    pk = packed_of(name)
    if name == "cfg1":
        assert str(g["digest"]) == synth.make_system(1).digest()
    for spin in (0, 1):
        got = cusp_fit(pk, spin)
        ref_rc = g["cusp_rc"]
        assert got.shape == ref_rc.shape
        np.testing.assert_array_equal(got["rc"] > 0, ref_rc > 0)
        on = ref_rc > 0
        for k in ("rc", "C", "sgn0", "phi0", "n0", "zeff", "sigma_sq", "ideal0"):
            a, b = got[k][on], g["cusp_" + k][on]
            err = np.abs(a - b) / np.maximum(1e-12, np.abs(b))
            ok = (err < 1e-9) | (np.abs(a - b) < 1e-12)
            assert ok.all(), (name, k, float(err.max()), int((~ok).sum()))
        a, b = got["alpha"][on], g["cusp_alpha"][on]
        assert np.all(np.abs(a - b) <= 1e-9 * np.maximum(1.0, np.abs(b))), name


def test_cusp_fit_covers_shifted_and_signed_cases():
    """The golden set exercises both signs of the exponential and the
    shifted form (C != 0) that needs the node search."""
    g = np.load(os.path.join(GOLDEN, "cusp_cfg1.npz"))
    on = g["cusp_rc"] > 0
    assert (g["cusp_sgn0"][on] > 0).any() and (g["cusp_sgn0"][on] < 0).any()
    assert (g["cusp_C"][on] != 0).any()


def test_cusp_replacement_matches_its_conditions():
    """Closed-form pins of the fitted quartic: at rc it matches value, slope
    and curvature of the original s part; at 0 its log-slope is the cusp
    value x4 and its value phi0."""
    pk = synth.make_system(1).pack()
    t = cusp_fit(pk, 0)
    nn, norb = t.shape
    B = int(pk["n_basis"])
    checked = 0
    for a in range(nn):
        # s functions on atom a
        js = [j for j in range(B) if pk["bf_atom"][j] == a and not pk["bf_pow"][j].any()]
        for m in range(0, norb, 7):
            f = t[a, m]
            if f["rc"] <= 0:
                continue
            al, cc, sg = f["alpha"], f["C"], f["sgn0"]
            rc = f["rc"]

            def orig(r):
                v = d1 = d2 = 0.0
                for j in js:
                    for p in range(pk["bf_prim_off"][j], pk["bf_prim_off"][j + 1]):
                        c = pk["prim_coef"][p] * pk["mo_up"][m, j]
                        e = pk["prim_alpha"][p]
                        tt = c * np.exp(-e * r * r)
                        v += tt
                        d1 += -2 * e * r * tt
                        d2 += (4 * e * e * r * r - 2 * e) * tt
                return v, d1, d2

            def rep(r):
                p = np.polynomial.polynomial.Polynomial(al)
                e = sg * np.exp(p(r))
                return cc + e, e * p.deriv()(r), e * (p.deriv(2)(r) + p.deriv()(r) ** 2)

            o, r_ = orig(rc), rep(rc)
            for u, v in zip(o, r_):
                assert abs(u - v) <= 1e-8 * max(1.0, abs(u)), (a, m, o, r_)
            assert abs(rep(0.0)[0] - f["phi0"]) <= 1e-10 * max(1.0, abs(f["phi0"]))
            checked += 1
    assert checked > 10
