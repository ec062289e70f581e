"""Pins of the C oracle (oracle/oracle.c) -- CPU only.

1. Golden vectors written by the reference's own evaluate compiled from its
   sources (tests/golden/*.npz, see make_golden.py): all-electron and
   single-electron Sherman-Morrison sequences, plus the umrigar93 drift.
2. What mathematics fixes independently of the reference: closed forms of an
   s Gaussian and of the Pade function, finite differences of ln|Psi| for the
   gradient and Laplacian, numpy slogdet/inv for the LU, brute-force pair sums
   for Coulomb and Jastrow, and the hydrogen-like local energy.
"""
import os

import numpy as np
import pytest

from qmcbeaver_amd import synth
from oracle import pyoracle
from oracle.pyoracle import Oracle

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")
CASES = [("cfg0_h2", 0), ("cfg1_organic64", 1), ("cfg2_mol256", 2)]


def rel(a, b):
    a, b = np.asarray(a, float), np.asarray(b, float)
    return float(np.max(np.abs(a - b) / np.maximum(1.0, np.abs(b))))


@pytest.fixture(scope="module", params=CASES, ids=[c[0] for c in CASES])
def golden(request):
    name, cfg = request.param
    g = np.load(os.path.join(GOLDEN, name + ".npz"))
    s = synth.make_system(cfg)
    assert s.digest() == str(g["digest"]), "generator drifted from the fixture"
    return s, g


def test_golden_all_electron(golden):
    s, g = golden
    o = Oracle(s.pack())
    res = o.evaluate_all(g["R"])
    for k in ("log_psi", "e_local", "kinetic", "potential", "v_ne", "v_ee"):
        assert rel(res[k], g["ev_" + k]) < 1e-10, k
    assert rel(res["grad_lnpsi"], g["ev_grad_lnpsi"]) < 1e-10
    assert rel(res["mod_drift"], g["ev_mod_drift"]) < 1e-10
    assert rel(res["mod_ratio"], g["ev_mod_ratio"]) < 1e-12
    np.testing.assert_allclose(res["r2"], (g["R"] ** 2).sum(axis=1), rtol=1e-14)
    assert not res["singular"].any()


def test_golden_umrigar93_drift(golden):
    s, g = golden
    if "umr_mod_drift" not in g:
        pytest.skip("no umrigar93 fixture for this size")
    o = Oracle(s.pack(), drift_mod=1, dt=0.1)
    n = g["umr_mod_drift"].shape[0]
    res = o.evaluate_all(g["R"][:n])
    assert rel(res["mod_drift"], g["umr_mod_drift"]) < 1e-10
    assert rel(res["e_local"], g["umr_e_local"]) < 1e-10
    assert rel(res["mod_ratio"], g["umr_mod_ratio"]) < 1e-10
    assert np.all(g["umr_mod_ratio"] < 1.0)  # the modification shortens the drift


def test_golden_single_electron_sequence(golden):
    s, g = golden
    o = Oracle(s.pack())
    mv = g["moves"]
    nw, nm = mv.shape[:2]
    j = 0
    for w in range(nw):
        wk = o.new_walker()
        Rw = g["R"][w].copy()
        r = o.evaluate(wk, Rw, -1)
        recs = [r]
        for m in range(nm):
            e = int(mv[w, m, 0])
            Rw[e] = mv[w, m, 1:]
            recs.append(o.evaluate(wk, Rw, e))
        for r in recs:
            # Ill-conditioned random Slater matrices: SM updates keep ~1e-10.
            assert rel(r["log_psi"], g["sq_log_psi"][j]) < 1e-9
            assert rel(r["e_local"], g["sq_e_local"][j]) < 1e-9
            assert rel(r["grad_lnpsi"][0], g["sq_grad_lnpsi"][j]) < 1e-9
            j += 1


# ----------------------------------------------------------- closed forms --
def one_center_system(shells, Z=1.0, n_up=1, n_dn=0, jastrow=False, mo=None):
    """Tiny system: one nucleus at the origin with the given shells."""
    sh = [synth.Shell(l, np.array(a, float), np.array(c, float)) for l, a, c in shells]
    B = sum(len(synth.CART_TYPES[x.l]) for x in sh)
    if mo is None:
        mo = np.eye(max(n_up, 1), B)
    zero = ([], [])
    s = synth.System(name="t", labels=["H"], Z=np.array([int(Z)]), xyz=np.zeros((1, 3)),
                     shells=[sh], mo=np.asarray(mo, float), n_up=n_up, n_dn=n_dn,
                     charge=int(Z) - n_up - n_dn,
                     ee_updn=([0.5], [1.0]) if jastrow else zero,
                     ee_upup=([0.25], [1.0]) if jastrow else zero,
                     en={"H": ([-1.0], [1.0]) if jastrow else zero})
    p = s.pack()
    p["nuc_z"] = np.array([Z])
    return s, p


def test_basis_s_gaussian_closed_form():
    al, c = 0.7, 0.9
    s, p = one_center_system([(0, [al], [c])])
    o = Oracle(p)
    r = np.array([0.3, -0.4, 0.5])
    v, gx, gy, gz, lap = o.basis(r)
    r2 = r @ r
    chi = c * np.exp(-al * r2)
    assert abs(v[0] - chi) < 1e-15
    np.testing.assert_allclose([gx[0], gy[0], gz[0]], -2 * al * r * chi, rtol=1e-13)
    np.testing.assert_allclose(lap[0], (4 * al * al * r2 - 6 * al) * chi, rtol=1e-13)


def test_basis_finite_differences_pd():
    shells = [(1, [0.8, 0.3], [0.5, 0.7]), (2, [1.1, 0.4, 0.2], [0.3, 0.6, 0.9]),
              (0, [2.0, 0.5], [0.2, 0.8])]
    s, p = one_center_system(shells)
    o = Oracle(p)
    r = np.array([0.37, -0.61, 0.44])
    v, gx, gy, gz, lap = o.basis(r)
    h = 1e-4
    g_fd = np.zeros((3, len(v)))
    l_fd = -6 * v
    for k in range(3):
        dp = r.copy(); dp[k] += h
        dm = r.copy(); dm[k] -= h
        vp, vm = o.basis(dp)[0], o.basis(dm)[0]
        g_fd[k] = (vp - vm) / (2 * h)
        l_fd = l_fd + vp + vm
    l_fd /= h * h
    np.testing.assert_allclose(np.array([gx, gy, gz]), g_fd, rtol=1e-7, atol=1e-9)
    np.testing.assert_allclose(lap, l_fd, rtol=1e-5, atol=1e-6)


def test_pade_closed_form():
    a, b = 0.5, 1.3
    row = synth._pade_arr(([a], [b]))
    for r in (0.05, 0.7, 3.0, 11.0):
        u, du, d2u = pyoracle.pade(row, r)
        assert abs(u - a * r / (1 + b * r)) < 1e-15
        assert abs(du - a / (1 + b * r) ** 2) < 1e-14
        assert abs(d2u + 2 * a * b / (1 + b * r) ** 3) < 1e-13
    row2 = synth._pade_arr(([0.3, -0.2], [0.9, 0.4]))
    r, h = 1.7, 1e-5
    u, du, d2u = pyoracle.pade(row2, r)
    up, _, _ = pyoracle.pade(row2, r + h)
    um, _, _ = pyoracle.pade(row2, r - h)
    assert abs(du - (up - um) / (2 * h)) < 1e-8
    assert abs(d2u - (up - 2 * u + um) / h ** 2) < 1e-4


def test_lu_inverse_matches_numpy():
    rng = np.random.default_rng(5)
    for n in (1, 2, 7, 32, 65):
        A = rng.normal(size=(n, n))
        ok, inv, det = pyoracle.det_inverse(A)
        assert ok
        sgn, ld = np.linalg.slogdet(A)
        assert np.sign(det) == sgn
        assert abs(np.log(abs(det)) - ld) < 1e-11 * max(1, abs(ld))
        np.testing.assert_allclose(inv @ A, np.eye(n), atol=1e-10)
    ok, _, _ = pyoracle.det_inverse(np.zeros((3, 3)))
    assert not ok


def test_sherman_morrison_row_update():
    rng = np.random.default_rng(6)
    n = 24
    A = rng.normal(size=(n, n))
    inv = np.linalg.inv(A)  # inv(o,e) for D(e,o) = A
    for row in (0, 5, n - 1):
        newrow = rng.normal(size=n)
        ratio, inv2 = pyoracle.inverse_update_row(inv, row, newrow)
        A2 = A.copy(); A2[row] = newrow
        np.testing.assert_allclose(inv2, np.linalg.inv(A2), rtol=1e-8, atol=1e-9)
        assert abs(ratio - np.linalg.det(A2) / np.linalg.det(A)) < 1e-9 * abs(ratio)


def test_hydrogen_like_local_energy():
    al, Z = 0.6, 1.0
    s, p = one_center_system([(0, [al], [1.0])], Z=Z)
    o = Oracle(p)
    R = np.array([[0.4, 0.2, -0.3]])
    res = o.evaluate_all(R[None])
    r = np.linalg.norm(R[0])
    el = -0.5 * (4 * al * al * r * r - 6 * al) - Z / r
    assert abs(res["e_local"][0] - el) < 1e-13
    np.testing.assert_allclose(res["grad_lnpsi"][0, 0], -2 * al * R[0], rtol=1e-13)


def _fd_logpsi(o, R, h=1e-4):
    N = R.shape[0]
    base = o.evaluate_all(R[None])
    lp0 = base["log_psi"][0]
    g = np.zeros((N, 3))
    lap = 0.0
    for e in range(N):
        for k in range(3):
            Rp = R.copy(); Rp[e, k] += h
            Rm = R.copy(); Rm[e, k] -= h
            lpp = o.evaluate_all(Rp[None])["log_psi"][0]
            lpm = o.evaluate_all(Rm[None])["log_psi"][0]
            d1 = (lpp - lpm) / (2 * h)
            d2 = (lpp - 2 * lp0 + lpm) / h ** 2
            g[e, k] = d1
            lap += d2 + d1 * d1
    return base, g, lap


def test_evaluate_finite_differences_small_molecule():
    shells = [(0, [1.2, 0.3], [0.6, 0.8]), (1, [0.9, 0.25], [0.5, 0.4]),
              (0, [0.15], [0.7])]
    mo = np.linalg.qr(np.random.default_rng(3).normal(size=(5, 2)))[0].T
    s, p = one_center_system(shells, Z=4.0, n_up=2, n_dn=2, jastrow=True,
                             mo=np.vstack([mo]))
    p["mo_dn"] = p["mo_up"].copy()
    o = Oracle(p)
    R = np.random.default_rng(4).normal(scale=0.8, size=(4, 3))
    base, g, lap = _fd_logpsi(o, R)
    np.testing.assert_allclose(base["grad_lnpsi"][0], g, rtol=1e-6, atol=1e-6)
    assert abs(base["lap_ratio"][0] - lap) < 2e-4 * max(1, abs(lap))


def test_coulomb_and_jastrow_brute_force():
    s = synth.make_system(1)
    o = Oracle(s.pack())
    R = synth.initial_walkers(s, 2, 7)
    res = o.evaluate_all(R)
    for w in range(2):
        X = R[w]
        N, na = s.n_elec, s.n_up
        d = np.linalg.norm(X[:, None] - X[None], axis=-1)
        iu = np.tril_indices(N, -1)
        vee = np.sum(1.0 / d[iu])
        dA = np.linalg.norm(X[:, None] - s.xyz[None], axis=-1)
        vne = -np.sum(s.Z[None] / dA)
        assert abs(res["v_ee"][w] - vee) < 1e-10 * abs(vee)
        assert abs(res["v_ne"][w] - vne) < 1e-10 * abs(vne)
        U = 0.0
        for i, j in zip(*iu):
            same = (i < na) == (j < na)
            a = 0.25 if same else 0.5
            U += a * d[i, j] / (1 + d[i, j])
        U += np.sum(-1.0 * dA / (1 + dA))
        assert abs(res["jastrow_u"][w] - U) < 1e-10 * abs(U)


@pytest.mark.parametrize("name,cfg", [("cfg0_h2", 0), ("cfg1_organic64", 1)])
def test_golden_umrigar93_equalelectrons_drift(name, cfg):
    """The equal-electron drift factor (one factor from |grad ln Psi|^2 of all
    electrons, umrigar93_equalelectrons_parameter 0.5, dt 0.1) against the
    reference's evaluate (tests/golden/make_golden_drift_equal.py)."""
    g = np.load(os.path.join(GOLDEN, "drift_equal.npz"))
    s = synth.make_system(cfg)
    assert s.digest() == str(g[name + "_digest"])
    o = Oracle(s.pack(), drift_mod=2, dt=0.1, umrigar_a=0.5)
    res = o.evaluate_all(g[name + "_R"])
    assert rel(res["grad_lnpsi"], g[name + "_grad_lnpsi"]) < 1e-10
    assert rel(res["mod_drift"], g[name + "_mod_drift"]) < 1e-10
    # the factor is one scalar per walker and below 1
    f = g[name + "_mod_drift"] / g[name + "_grad_lnpsi"]
    assert np.allclose(f, f[:, :1, :1], rtol=1e-9) and (f < 1).all()
