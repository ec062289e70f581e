"""Write the golden vectors of the DDDA statistics (ddda_cases.npz).

Runs the reference's own QMCProperty (QMCProperty.cpp, QMCStatistic.cpp),
linked into oracle/_ref/qmcb_ref_driver by oracle/ref.mk, on seeded weighted
series, and stores the series and the reference's getters.  Run in the
authoring container only (needs /root/reference):

    make -f oracle/ref.mk && python tests/golden/make_golden_stats.py
"""
from __future__ import annotations

import json
import os
import subprocess
import sys
import tempfile

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
DRIVER = os.path.join(ROOT, "oracle", "_ref", "qmcb_ref_driver")
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "ddda_cases.npz")


def ar1(rng, n, rho, mean=0.0, sigma=1.0):
    x = np.empty(n)
    x[0] = rng.normal() * sigma / np.sqrt(1 - rho * rho)
    e = rng.normal(size=n) * sigma
    for i in range(1, n):
        x[i] = rho * x[i - 1] + e[i]
    return x + mean


def cases():
    rng = np.random.default_rng(0x51EED05)
    # name, series, weights, split (split < n: two accumulations merged)
    a = ar1(rng, 2000, 0.8)
    yield "ar1", a, np.ones_like(a), 2000
    yield "ar1_split", a, np.ones_like(a), 777
    x = rng.normal(size=2000) * 0.3 - 1.1
    w = rng.uniform(0.3, 3.0, size=2000)
    yield "weighted", x, w, 2000
    yield "weighted_split", x, w, 1024
    s = rng.normal(size=60)
    yield "short", s, np.ones_like(s), 60
    d = ar1(rng, 2048, 0.9, mean=-1.17, sigma=0.01)
    dw = 1000.0 + np.cumsum(rng.normal(size=2048))
    yield "dmc_like", d, dw, 1000
    c = ar1(rng, 300, 0.999)
    yield "no_plateau", c, np.ones_like(c), 300


def run(x, w, split, tmp):
    f = os.path.join(tmp, "s.bin")
    with open(f, "wb") as fh:
        np.array([x.size], dtype=np.int32).tofile(fh)
        np.stack([x, w], axis=1).astype(np.float64).tofile(fh)
    out = subprocess.run([DRIVER, "ddda", f, str(split)], check=True, capture_output=True,
                         text=True).stdout
    return json.loads(out)


def nan(v):
    return np.nan if v is None else float(v)


def main():
    if not os.path.exists(DRIVER):
        sys.exit("build the reference driver first: make -f oracle/ref.mk")
    data = {}
    names = []
    stored = {}
    with tempfile.TemporaryDirectory() as tmp:
        for name, x, w, split in cases():
            r = run(x, w, split, tmp)
            names.append(name)
            key = (x.ctypes.data, w.ctypes.data)
            if key not in stored:  # a series shared by two cases is stored once
                stored[key] = name
                data[name + "_x"] = x
                data[name + "_w"] = w
            data[name + "_src"] = stored[key]
            data[name + "_split"] = split
            data[name + "_ref"] = np.array([r["n"], nan(r["average"]), nan(r["std"]),
                                            nan(r["corr_len"]), nan(r["serial_var"]),
                                            nan(r["std_std"])])
            data[name + "_block_var"] = np.array([nan(v) for v in r["block_var"]])
            print(name, r["average"], r["std"], r["corr_len"])
    np.savez_compressed(OUT, names=np.array(names), **data)


if __name__ == "__main__":
    main()
