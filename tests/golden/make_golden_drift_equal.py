"""Golden vectors of the umrigar93_equalelectrons drift (one factor for all
electrons, QMCSCFJastrow.cpp:340-366): all-electron evaluates of the 2- and
64-electron synthetic systems by the reference's own evaluate
(oracle/_ref/qmcb_ref_driver) with QF_modification_type
umrigar93_equalelectrons, umrigar93_equalelectrons_parameter 0.5 and dt 0.1.
Authoring container only (needs /root/reference):

    make -f oracle/ref.mk && python tests/golden/make_golden_drift_equal.py
"""
from __future__ import annotations

import os
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
sys.path.insert(0, HERE)
from qmcbeaver_amd import synth  # noqa: E402
from make_golden import DRIVER, parse  # noqa: E402

CASES = [("cfg0_h2", 0, 16, 201), ("cfg1_organic64", 1, 6, 202)]


def run_eval_equal(sys_, R, tmp):
    import subprocess
    ck = os.path.join(tmp, "in.ckmf")
    synth.write_ckmf(sys_, ck, dt=0.1,
                     extra_flags={"QF_modification_type": "umrigar93_equalelectrons",
                                  "umrigar93_equalelectrons_parameter": "0.5"})
    cf = os.path.join(tmp, "cfg.bin")
    with open(cf, "wb") as f:
        np.array([R.shape[0], R.shape[1]], dtype=np.int32).tofile(f)
        np.ascontiguousarray(R, dtype=np.float64).tofile(f)
    of = os.path.join(tmp, "out.bin")
    subprocess.run([DRIVER, "eval", ck, cf, of], check=True, cwd=tmp,
                   stdout=subprocess.DEVNULL)
    return parse(np.fromfile(of, dtype=np.float64), R.shape[1])


def main():
    if not os.path.exists(DRIVER):
        sys.exit("build the reference driver first: make -f oracle/ref.mk")
    out = {}
    with tempfile.TemporaryDirectory() as tmp:
        for name, cfg, nw, seed in CASES:
            s = synth.make_system(cfg)
            R = synth.initial_walkers(s, nw, seed)
            ev = run_eval_equal(s, R, tmp)
            out[name + "_R"] = R
            out[name + "_digest"] = s.digest()
            for k in ("mod_drift", "grad_lnpsi", "e_local", "log_psi"):
                out[name + "_" + k] = ev[k]
            print(name, "mod_drift[0,0] =", ev["mod_drift"][0, 0], "grad[0,0] =", ev["grad_lnpsi"][0, 0])
    np.savez_compressed(os.path.join(HERE, "drift_equal.npz"), **out)


if __name__ == "__main__":
    main()
