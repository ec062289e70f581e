"""Helpers for the reference-input fixtures (tests/golden/*.ckmf.gz)."""
import gzip
import os
import tempfile

import numpy as np

from qmcbeaver_amd import ckmf

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
INPUTS = ["h2_631gss", "hmx"]
OVERRIDES = {"replace_electron_nucleus_cusps": "0", "one_e_per_iter": "0",
             "walkers_per_pass": "1", "QF_modification_type": "none",
             "use_available_checkpoints": "0"}


def load(name):
    """(packed system, golden npz) of a reference input with the overrides
    the golden vectors were written with."""
    with tempfile.TemporaryDirectory() as tmp:
        raw = os.path.join(tmp, name + ".ckmf")
        with gzip.open(os.path.join(GOLDEN, name + ".ckmf.gz"), "rb") as g, open(raw, "wb") as f:
            f.write(g.read())
        ck = os.path.join(tmp, "in.ckmf")
        ckmf.write_with_flags(raw, ck, OVERRIDES)
        packed, flags, info = ckmf.read_ckmf(ck)
    return packed, np.load(os.path.join(GOLDEN, name + ".npz")), flags
