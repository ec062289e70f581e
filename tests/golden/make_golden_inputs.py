"""Golden vectors for reference input files (the reference's own example and
test inputs, kept here gzipped as data): each file is read by the reference
build (oracle/_ref/qmcb_ref_driver) with cusp replacement switched off and
all-electron evaluation, at seeded configurations.  Run in the authoring
container only (needs /root/reference):

    make -C oracle ref && python tests/golden/make_golden_inputs.py
"""
from __future__ import annotations

import gzip
import os
import shutil
import subprocess
import sys
import tempfile

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
from qmcbeaver_amd import ckmf  # noqa: E402
from tests.golden.make_golden import DRIVER, parse  # noqa: E402
from tests.inputs_util import configs  # noqa: E402

OUT = os.path.dirname(os.path.abspath(__file__))
REF = "/root/reference/QMcBeaver"
INPUTS = [
    # fixture name, reference file, configurations
    ("h2_631gss", "examples/H2_hf_631Gss-vmc.ckmf", 32),
    ("hmx", "test/hmx.ckmf", 4),
]
OVERRIDES = {"replace_electron_nucleus_cusps": "0", "one_e_per_iter": "0",
             "walkers_per_pass": "1", "QF_modification_type": "none",
             "use_available_checkpoints": "0"}


def main():
    with tempfile.TemporaryDirectory() as tmp:
        for name, rel, n in INPUTS:
            src = os.path.join(REF, rel)
            gz = os.path.join(OUT, name + ".ckmf.gz")
            with open(src, "rb") as f, gzip.open(gz, "wb") as g:
                shutil.copyfileobj(f, g)
            results = {}
            for drift in ("none", "umrigar93_unequalelectrons"):
                ov = dict(OVERRIDES, QF_modification_type=drift)
                ck = os.path.join(tmp, name + ".ckmf")
                ckmf.write_with_flags(src, ck, ov)
                packed, flags, _ = ckmf.read_ckmf(ck)
                R = configs(packed, n, 777)
                cf = os.path.join(tmp, "cfg.bin")
                with open(cf, "wb") as f:
                    np.array(R.shape[:2], dtype=np.int32).tofile(f)
                    R.tofile(f)
                of = os.path.join(tmp, "out.bin")
                subprocess.run([DRIVER, "eval", ck, cf, of], check=True, cwd=tmp,
                               stdout=subprocess.DEVNULL)
                res = parse(np.fromfile(of, dtype=np.float64), R.shape[1])
                tag = "ev_" if drift == "none" else "umr_"
                results.update({tag + k: v for k, v in res.items()})
            np.savez_compressed(os.path.join(OUT, name + ".npz"), R=R, source=rel,
                                dt=ckmf.evaluate_dt(flags), **results)
            print(name, "N =", R.shape[1], "E_L[0] =", results["ev_e_local"][0])


if __name__ == "__main__":
    main()
