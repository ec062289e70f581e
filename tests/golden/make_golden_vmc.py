"""VMC energies of the reference's H2 6-31G** example
(examples/H2_hf_631Gss-vmc.ckmf, kept here as tests/golden/h2_631gss.ckmf.gz)
and of its 10-determinant Be MCSCF example (examples/BeMCSCF.ckmf, with the
reference's cusp replacement) with the reference evaluate and the reference
DDDA estimator, for the statistical parity tests of the device VMC.  Runs oracle/_ref/qmcb_ref_driver in its vmc mode: importance-sampled
single-electron moves with the Green's function ratio, rejection by
re-evaluation, one statistics sample per electron sweep from
p E_L(R') + (1 - p) E_L(R) averaged over walkers, fed to QMCProperty.
Authoring container only (needs /root/reference):

    make -f oracle/ref.mk && python tests/golden/make_golden_vmc.py [case ...]
"""
from __future__ import annotations

import gzip
import json
import os
import subprocess
import sys
import tempfile

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
from qmcbeaver_amd import ckmf  # noqa: E402
from tests.inputs_util import configs  # noqa: E402

DRIVER = os.path.join(ROOT, "oracle", "_ref", "qmcb_ref_driver")
OUT = os.path.dirname(os.path.abspath(__file__))
OVERRIDES = {"replace_electron_nucleus_cusps": "0", "one_e_per_iter": "1",
             "walkers_per_pass": "1", "QF_modification_type": "none",
             "use_available_checkpoints": "0", "optimize_Psi": "0"}
CASES = [
    ("h2_631gss", dict(walkers=2048, sweeps=6000, tau=0.05, seed=2024, config_seed=778), {}),
    ("be_mcscf", dict(walkers=1024, sweeps=3000, tau=0.02, seed=2025, config_seed=779),
     {"replace_electron_nucleus_cusps": "1"}),
    # the reference's default drift (umrigar93_unequalelectrons, time step
    # dt = tau) at a large time step, where the Green's function ratio
    # decides the acceptance; the reference starts a run with dt set to
    # dt_equilibration (QMCFlags.cpp:916), so both are set
    ("h2_631gss", dict(walkers=2048, sweeps=4000, tau=0.3, seed=2026, config_seed=780),
     {"QF_modification_type": "umrigar93_unequalelectrons", "dt": "0.3",
      "dt_equilibration": "0.3"}, "h2_631gss_qf"),
]


def main():
    if not os.path.exists(DRIVER):
        sys.exit("build the reference driver first: make -f oracle/ref.mk")
    only = set(sys.argv[1:])  # optional: names of the cases to (re)write
    for case in CASES:
        name, run, extra = case[:3]
        tag = case[3] if len(case) > 3 else name
        if not only or tag in only:
            one(name, run, dict(OVERRIDES, **extra), tag)


def one(name, RUN, overrides, tag):
    with tempfile.TemporaryDirectory() as tmp:
        raw = os.path.join(tmp, "in_raw.ckmf")
        with gzip.open(os.path.join(OUT, name + ".ckmf.gz"), "rb") as g, open(raw, "wb") as f:
            f.write(g.read())
        ck = os.path.join(tmp, "in.ckmf")
        ckmf.write_with_flags(raw, ck, overrides)
        packed, flags, _ = ckmf.read_ckmf(ck)
        R = configs(packed, RUN["walkers"], RUN["config_seed"])
        cf = os.path.join(tmp, "cfg.bin")
        with open(cf, "wb") as f:
            np.array(R.shape[:2], dtype=np.int32).tofile(f)
            R.tofile(f)
        N = R.shape[1]
        out = subprocess.run([DRIVER, "vmc", ck, cf, str(RUN["sweeps"] * N), str(RUN["tau"]),
                              str(RUN["seed"])], check=True, capture_output=True, text=True,
                             cwd=tmp).stdout
        res = json.loads(out.strip().splitlines()[-1])
    res.pop("seconds", None)
    rec = dict(RUN, input=name, overrides=overrides, reference=res)
    with open(os.path.join(OUT, f"vmc_{tag}.json"), "w") as f:
        json.dump(rec, f, indent=1)
    print(rec)


if __name__ == "__main__":
    main()
