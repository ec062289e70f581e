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


def configs(packed, n, seed):
    """n configurations: electron k of each spin near nucleus order[2k (+1)]
    (nuclei repeated Z times), displaced by N(0, 0.5^2)."""
    rng = np.random.default_rng(seed)
    Z = np.maximum(packed["nuc_z"].astype(int), 1)
    order = np.repeat(np.arange(packed["n_nuc"]), Z)
    nuc = np.array([order[(2 * k) % len(order)] for k in range(packed["n_up"])] +
                   [order[(2 * k + 1) % len(order)] for k in range(packed["n_dn"])])
    N = packed["n_up"] + packed["n_dn"]
    return packed["nuc_xyz"][nuc][None] + 0.5 * rng.normal(size=(n, N, 3))
