"""The QMCFunctions-shaped C++ interface (include/qmcb_functions.hpp) driving
a host VMC pass over the C ABI, as QMCRun::propagateWalkers drives
QMCFunctions::evaluate in the reference."""
import json
import os
import subprocess
import tempfile

import numpy as np
import pytest

from qmcbeaver_amd import synth, sysio, Context

pytestmark = pytest.mark.gpu
EXE = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                   "qmcbeaver_amd", "examples", "dropin_vmc")


def test_cpp_dropin_matches_python_binding():
    s = synth.make_system(1)
    R = synth.initial_walkers(s, 16, 31)
    with tempfile.TemporaryDirectory() as tmp:
        sysio.write_flat(s.pack(), os.path.join(tmp, "s.bin"))
        sysio.write_walkers(R, os.path.join(tmp, "w.bin"))
        print("case cfg1 C++ drop-in, 16 walkers x 64 moves", flush=True)
        out = subprocess.run([EXE, os.path.join(tmp, "s.bin"), os.path.join(tmp, "w.bin"), "64"],
                             capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    r = json.loads(out.stdout.strip().splitlines()[-1])
    with Context(s.pack(), precision="fp64", max_walkers=16) as ctx:
        ref = ctx.evaluate(np.arange(16), R)
    assert abs(r["e_local_0"] - ref["e_local"][0]) <= 1e-12 * max(1, abs(ref["e_local"][0]))
    assert 0.2 < r["acceptance"] <= 1.0
    assert np.isfinite(r["mean_prop_e_local"])
