"""Write the golden vectors of the walker-evaluation path.

Runs the reference's own evaluate (oracle/_ref/qmcb_ref_driver, compiled from
the reference sources by oracle/ref.mk) on seeded synthetic systems and
stores inputs (seeds, system digest, configurations/moves) and outputs as
small .npz fixtures beside this script.  Run in the authoring container only
(needs /root/reference):

    make -C oracle ref && python tests/golden/make_golden.py [case ...]

(no case names: every case)
"""
from __future__ import annotations

import os
import subprocess
import sys
import tempfile

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
from qmcbeaver_amd import synth  # noqa: E402

DRIVER = os.path.join(ROOT, "oracle", "_ref", "qmcb_ref_driver")
OUT = os.path.dirname(os.path.abspath(__file__))
NFIX = 12


def parse(raw: np.ndarray, N: int):
    rec = raw.reshape(-1, NFIX + 6 * N)
    return {
        "detA": rec[:, 0], "detB": rec[:, 1], "U": rec[:, 2], "e_local": rec[:, 3],
        "kinetic": rec[:, 4], "potential": rec[:, 5], "v_ne": rec[:, 6],
        "v_ee": rec[:, 7], "singular": rec[:, 8], "mod_ratio": rec[:, 9],
        "log_psi": rec[:, 10], "sign_psi": rec[:, 11],
        "grad_lnpsi": rec[:, NFIX:NFIX + 3 * N].reshape(-1, N, 3),
        "mod_drift": rec[:, NFIX + 3 * N:].reshape(-1, N, 3),
    }


def run_eval(sys_, R, tmp, drift="none", dt=0.1):
    ck = os.path.join(tmp, "in.ckmf")
    synth.write_ckmf(sys_, ck, dt=dt, extra_flags={"QF_modification_type": drift})
    cf = os.path.join(tmp, "cfg.bin")
    with open(cf, "wb") as f:
        np.array([R.shape[0], R.shape[1]], dtype=np.int32).tofile(f)
        np.ascontiguousarray(R, dtype=np.float64).tofile(f)
    of = os.path.join(tmp, "out.bin")
    subprocess.run([DRIVER, "eval", ck, cf, of], check=True, cwd=tmp,
                   stdout=subprocess.DEVNULL)
    return parse(np.fromfile(of, dtype=np.float64), R.shape[1])


def run_seq(sys_, R0, moves, tmp):
    """moves: (nw, nmoves, 4) = electron index, new x, y, z."""
    ck = os.path.join(tmp, "in.ckmf")
    synth.write_ckmf(sys_, ck)
    cf = os.path.join(tmp, "mv.bin")
    nw, nm = moves.shape[:2]
    with open(cf, "wb") as f:
        np.array([nw, R0.shape[1], nm], dtype=np.int32).tofile(f)
        for w in range(nw):
            np.ascontiguousarray(R0[w], dtype=np.float64).tofile(f)
            np.ascontiguousarray(moves[w], dtype=np.float64).tofile(f)
    of = os.path.join(tmp, "out.bin")
    subprocess.run([DRIVER, "seq", ck, cf, of], check=True, cwd=tmp,
                   stdout=subprocess.DEVNULL)
    return parse(np.fromfile(of, dtype=np.float64), R0.shape[1])


def make_moves(sys_, R0, nmoves, seed, step=0.3):
    """Single-electron displacement sequence: electron m mod N moves by
    N(0, step^2) in every walker."""
    rng = np.random.default_rng(seed)
    nw, N = R0.shape[:2]
    mv = np.zeros((nw, nmoves, 4))
    R = R0.copy()
    for m in range(nmoves):
        e = m % N
        R[:, e, :] += step * rng.normal(size=(nw, 3))
        mv[:, m, 0] = e
        mv[:, m, 1:] = R[:, e, :]
    return mv


CASES = [
    # name, config, n_walkers (eval), walker seed, n_seq_walkers, n_moves
    ("cfg0_h2", 0, 128, 101, 8, 40),
    ("cfg1_organic64", 1, 24, 102, 2, 70),
    ("cfg2_mol256", 2, 4, 103, 1, 136),  # moves 128 up and 8 down electrons
]


def main():
    if not os.path.exists(DRIVER):
        sys.exit("build the reference driver first: make -C oracle ref")
    only = set(sys.argv[1:])
    with tempfile.TemporaryDirectory() as tmp:
        for name, cfg, nw, wseed, nsw, nmov in CASES:
            if only and name not in only:
                continue
            s = synth.make_system(cfg)
            R = synth.initial_walkers(s, nw, wseed)
            ev = run_eval(s, R, tmp)
            extra = {}
            if cfg in (0, 1, 2):
                ev_u = run_eval(s, R[: min(nw, 8)], tmp, drift="umrigar93_unequalelectrons", dt=0.1)
                extra = {"umr_" + k: v for k, v in ev_u.items()}
            R0 = R[:nsw]
            mv = make_moves(s, R0, nmov, wseed + 1000)
            sq = run_seq(s, R0, mv, tmp)
            np.savez_compressed(
                os.path.join(OUT, name + ".npz"),
                config=cfg, walker_seed=wseed, n_walkers=nw, digest=s.digest(),
                R=R, moves=mv, **{"ev_" + k: v for k, v in ev.items()},
                **{"sq_" + k: v for k, v in sq.items() if k != "mod_drift"}, **extra)
            print(name, "written; E_L[0] =", ev["e_local"][0])


if __name__ == "__main__":
    main()
