"""Golden vectors of the electron-nucleus cusp replacement (cusp_*.npz).

The reference build (oracle/_ref/qmcb_ref_driver, reference sources compiled
by oracle/ref.mk) reads each input with replace_electron_nucleus_cusps 1 and
print_replacement_orbitals 1: it prints every fitted replacement orbital
(QMCElectronNucleusCuspParameters::printParameters, at 17 digits) and then
evaluates seeded configurations, some with electrons placed inside the
correction radii, all-electron and as a single-electron move sequence.
Authoring container only (needs /root/reference):

    make -f oracle/ref.mk && python tests/golden/make_golden_cusp.py
"""
from __future__ import annotations

import gzip
import os
import re
import subprocess
import sys
import tempfile

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
from qmcbeaver_amd import ckmf, synth  # noqa: E402
from tests.inputs_util import configs  # noqa: E402

DRIVER = os.path.join(ROOT, "oracle", "_ref", "qmcb_ref_driver")
OUT = os.path.dirname(os.path.abspath(__file__))
NFIX = 12
FLAGS = {"replace_electron_nucleus_cusps": "1", "print_replacement_orbitals": "1",
         "walkers_per_pass": "1", "QF_modification_type": "none",
         "use_available_checkpoints": "0"}
FIELDS = ("rc", "C", "sgn0", "phi0", "n0", "zeff", "sigma_sq", "ideal0")


def parse_eval(raw, N):
    rec = raw.reshape(-1, NFIX + 6 * N)
    return {"detA": rec[:, 0], "detB": rec[:, 1], "U": rec[:, 2], "e_local": rec[:, 3],
            "kinetic": rec[:, 4], "potential": rec[:, 5], "v_ne": rec[:, 6], "v_ee": rec[:, 7],
            "singular": rec[:, 8], "log_psi": rec[:, 10], "sign_psi": rec[:, 11],
            "grad_lnpsi": rec[:, NFIX:NFIX + 3 * N].reshape(-1, N, 3)}


def parse_params(text, nn, norb):
    """Fitted tables in print order (the reference fits once per spin and
    Slater object); returns the first and checks that every other one is
    identical (restricted inputs)."""
    tabs = []
    cur = None
    for blk in re.split(r"Orbital replacement parameters for atom ", text)[1:]:
        m = re.match(r"(\d+) orbital (\d+)", blk)
        a, o = int(m.group(1)), int(m.group(2))
        if cur is None or (a, o) in cur["seen"]:
            cur = {k: np.zeros((nn, norb)) for k in FIELDS}
            cur["alpha"] = np.zeros((nn, norb, 5))
            cur["seen"] = set()
            tabs.append(cur)
        cur["seen"].add((a, o))

        def num(key):
            return float(re.search(re.escape(key) + r" = (\S+)", blk).group(1))
        cur["rc"][a, o] = num("rc")
        cur["C"][a, o] = num("C")
        cur["sgn0"][a, o] = num("sgn0")
        cur["phi0"][a, o] = num("phi0")
        cur["n0"][a, o] = num("n0")
        cur["zeff"][a, o] = num("Zeff")
        cur["sigma_sq"][a, o] = num("sigmaSq")
        al = re.search(r"alpha coefficients:\s*\n\s*(.+)\n", blk).group(1).split()
        cur["alpha"][a, o] = [float(v) for v in al]
        ic = re.search(r"ideal curve coefficients:\s*\n\s*(.+)\n", blk).group(1).split()
        cur["ideal0"][a, o] = float(ic[0])
    assert len(tabs) >= 2, len(tabs)
    for t in tabs[1:]:
        for k in FIELDS + ("alpha",):
            np.testing.assert_array_equal(tabs[0][k], t[k])
    return {k: tabs[0][k] for k in FIELDS + ("alpha",)}


def near_nucleus_configs(packed, n, seed):
    """configs() with, in every configuration, a few electrons of each spin
    moved to 0.02-0.9 of 1/Z from a nucleus (inside typical radii)."""
    R = configs(packed, n, seed)
    rng = np.random.default_rng(seed + 1)
    nu, N, nn = int(packed["n_up"]), R.shape[1], int(packed["n_nuc"])
    for c in range(n):
        picks = list(rng.choice(nu, size=min(3, nu), replace=False))
        if N > nu:
            picks += list(nu + rng.choice(N - nu, size=min(3, N - nu), replace=False))
        for e in picks:
            a = int(rng.integers(nn))
            d = rng.normal(size=3)
            d *= rng.uniform(0.02, 0.9) / max(1.0, float(packed["nuc_z"][a])) / np.linalg.norm(d)
            R[c, e] = packed["nuc_xyz"][a] + d
    return R


def run(ck, R, moves, tmp):
    N = R.shape[1]
    cf = os.path.join(tmp, "cfg.bin")
    with open(cf, "wb") as f:
        np.array(R.shape[:2], dtype=np.int32).tofile(f)
        R.tofile(f)
    of = os.path.join(tmp, "out.bin")
    r = subprocess.run([DRIVER, "eval", ck, cf, of], check=True, cwd=tmp, capture_output=True,
                       text=True)
    ev = parse_eval(np.fromfile(of, dtype=np.float64), N)
    if moves.shape[0] == 0:
        return r.stderr, ev, {}
    mf = os.path.join(tmp, "mv.bin")
    nw, nm = moves.shape[:2]
    with open(mf, "wb") as f:
        np.array([nw, N, nm], dtype=np.int32).tofile(f)
        for w in range(nw):
            R[w].tofile(f)
            moves[w].tofile(f)
    subprocess.run([DRIVER, "seq", ck.replace(".ckmf", "_1e.ckmf"), mf, of], check=True, cwd=tmp,
                   capture_output=True)
    sq = parse_eval(np.fromfile(of, dtype=np.float64), N)
    return r.stderr, ev, sq


def moves_into_cusps(packed, R0, nm, seed):
    rng = np.random.default_rng(seed)
    nw, N = R0.shape[:2]
    mv = np.zeros((nw, nm, 4))
    nn = int(packed["n_nuc"])
    for w in range(nw):
        for m in range(nm):
            e = (m * 7) % N
            if m % 2 == 0:  # into a correction region
                a = int(rng.integers(nn))
                d = rng.normal(size=3)
                d *= rng.uniform(0.02, 0.9) / max(1.0, float(packed["nuc_z"][a])) / np.linalg.norm(d)
                x = packed["nuc_xyz"][a] + d
            else:
                x = R0[w, e] + 0.3 * rng.normal(size=3)
            mv[w, m] = [e, *x]
    return mv


def main():
    if not os.path.exists(DRIVER):
        sys.exit("build the reference driver first: make -f oracle/ref.mk")
    # name, configurations, sequence walkers, sequence moves.  HMX has no
    # move sequence: its input lists more orbitals than electrons, where the
    # reference's one_e_per_iter row copy (Array2D::setRows, Array2D.h:964-987,
    # used at QMCSlater.cpp:486-490) reads with the wrong stride.
    cases = [("h2_631gss", 32, 2, 24), ("hmx", 4, 0, 0), ("cfg1", 8, 2, 30)]
    with tempfile.TemporaryDirectory() as tmp:
        for name, n, nsw, nmv in cases:
            ck = os.path.join(tmp, "in.ckmf")
            if name == "cfg1":
                s = synth.make_system(1)
                synth.write_ckmf(s, ck, extra_flags=dict(FLAGS, one_e_per_iter="0"))
                synth.write_ckmf(s, ck.replace(".ckmf", "_1e.ckmf"),
                                 extra_flags=dict(FLAGS, one_e_per_iter="1"))
                extra = {"digest": s.digest()}
            else:
                raw = os.path.join(tmp, name + ".ckmf")
                with gzip.open(os.path.join(OUT, name + ".ckmf.gz"), "rb") as g, \
                        open(raw, "wb") as f:
                    f.write(g.read())
                ckmf.write_with_flags(raw, ck, dict(FLAGS, one_e_per_iter="0"))
                ckmf.write_with_flags(raw, ck.replace(".ckmf", "_1e.ckmf"),
                                      dict(FLAGS, one_e_per_iter="1"))
                extra = {}
            packed, flags, _ = ckmf.read_ckmf(ck)
            R = near_nucleus_configs(packed, n, 4242)
            mv = moves_into_cusps(packed, R[:nsw], nmv, 4343)
            err, ev, sq = run(ck, R, mv, tmp)
            prm = parse_params(err, int(packed["n_nuc"]), int(packed["n_up"]))
            np.savez_compressed(
                os.path.join(OUT, f"cusp_{name}.npz"), R=R, moves=mv,
                **{"cusp_" + k: v for k, v in prm.items()},
                **{"ev_" + k: v for k, v in ev.items()}, **{"sq_" + k: v for k, v in sq.items()},
                **extra)
            print(name, "fitted", int((prm["rc"] > 0).sum()), "replacements; E_L[0]",
                  ev["e_local"][0])


if __name__ == "__main__":
    main()
