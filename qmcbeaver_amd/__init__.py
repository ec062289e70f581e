"""MI355X-native walker-evaluation path of QMcBeaver (QMCFunctions drop-in).

    from qmcbeaver_amd import synth, Context
    sys = synth.make_system(2)
    ctx = Context(sys.pack(), precision="fp32", max_walkers=8192)
"""
from .core import Context, QmcbError, lib, LIB_PATH, EXPORTS  # noqa: F401

__all__ = ["Context", "QmcbError", "lib", "LIB_PATH", "EXPORTS"]
