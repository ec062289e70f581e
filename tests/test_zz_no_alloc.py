"""qmcb_evaluate_device_noalloc (all-electron evaluate whose basis scratch is
a static device array instead of context memory) against
qmcb_evaluate_device on the same device inputs, through the C ABI.

Each walker's outputs depend only on its positions, and the chunking of the
points changes no accumulation order (the orbital GEMM sums each output over
the same k sequence whatever the tile), so the two entry points must agree
bit for bit.  Two calls in a row on the 256-electron FP64 system (the second
smaller), then a 64-electron FP32 system whose padded basis rows sit where
the first system's values were left in the shared scratch, then the first
system again.
"""
import ctypes as C
import sys

import numpy as np
import pytest

from qmcbeaver_amd import synth, Context, lib
from qmcbeaver_amd._abi import Results

pytestmark = pytest.mark.gpu

FIELDS = [("log_psi", "f8", 1), ("sign_psi", "i1", 1), ("grad_lnpsi", "f8", 3),
          ("mod_drift", "f8", 3), ("lap_ratio", "f8", 1), ("e_local", "f8", 1),
          ("kinetic", "f8", 1), ("potential", "f8", 1), ("v_ne", "f8", 1), ("v_ee", "f8", 1),
          ("det_ratio", "f8", 1), ("jastrow_u", "f8", 1), ("singular", "i1", 1),
          ("mod_ratio", "f8", 1), ("r2", "f8", 0)]


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


class Hip:
    """The HIP runtime the library links (libamdhip64.so.7, already loaded by
    libqmcb_amd.so, so this is the same runtime instance): device buffers for
    the inputs and outputs the caller owns."""

    H2D, D2H = 1, 2

    def __init__(self):
        lib()
        self.L = C.CDLL("libamdhip64.so.7")
        self.L.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.L.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.L.hipFree.argtypes = [C.c_void_p]
        self.bufs = []

    def alloc(self, nbytes):
        p = C.c_void_p()
        assert self.L.hipMalloc(C.byref(p), max(int(nbytes), 8)) == 0
        self.bufs.append(p)
        return p

    def upload(self, a):
        a = np.ascontiguousarray(a)
        p = self.alloc(a.nbytes)
        assert self.L.hipMemcpy(p, a.ctypes.data, a.nbytes, self.H2D) == 0
        return p

    def download(self, p, shape, dtype):
        a = np.empty(shape, dtype=dtype)
        assert self.L.hipMemcpy(a.ctypes.data, p, a.nbytes, self.D2H) == 0
        return a

    def free(self):
        for p in self.bufs:
            self.L.hipFree(p)
        self.bufs = []


def call(hip, ctx, entry, ids, R):
    L = lib()
    num, N = len(ids), R.shape[1]
    d_ids = hip.upload(np.asarray(ids, np.int32))
    d_R = hip.upload(np.asarray(R, np.float64))
    r, shapes = Results(), {}
    for name, kind, per in FIELDS:
        shape = (num, N, 3) if per == 3 else ((num, 3) if per == 0 else (num,))
        dt = np.float64 if kind == "f8" else np.int8
        p = hip.upload(np.full(shape, 7, dtype=dt))  # poisoned: every entry must be written
        shapes[name] = (p, shape, dt)
        setattr(r, name, C.cast(p, C.POINTER(C.c_double if kind == "f8" else C.c_int8)))
    if entry == "noalloc":
        st = L.qmcb_evaluate_device_noalloc(ctx._h, num, d_ids, d_R, C.byref(r))
    else:
        st = L.qmcb_evaluate_device(ctx._h, num, d_ids, d_R, -1, C.byref(r))
    assert st == 0, L.qmcb_last_error(ctx._h).decode()
    assert L.qmcb_sync(ctx._h) == 0
    out = {k: hip.download(p, shape, dt) for k, (p, shape, dt) in shapes.items()}
    hip.free()
    return out


def check_equal(a, b, what):
    for name, _, _ in FIELDS:
        assert np.array_equal(a[name], b[name], equal_nan=True), (what, name)
    assert not a["singular"].any(), what
    assert np.all(np.isfinite(a["e_local"])), what


def test_noalloc_evaluate_equals_evaluate_device():
    hip = Hip()
    s2, s1 = synth.make_system(2), synth.make_system(1)
    W2, W1 = 24, 40
    R2 = synth.initial_walkers(s2, W2, 71)
    R1 = synth.initial_walkers(s1, W1, 72)
    with Context(s2.pack(), precision="fp64", max_walkers=W2) as c2, \
            Context(s1.pack(), precision="fp32", max_walkers=W1) as c1:
        for num, seed in [(W2, 0), (7, 1)]:
            say("noalloc mol256 fp64", num)
            ids = np.random.default_rng(seed).permutation(W2)[:num]
            ref = call(hip, c2, "device", ids, R2[ids])
            got = call(hip, c2, "noalloc", ids, R2[ids])
            check_equal(got, ref, f"mol256 fp64 num={num}")
        say("noalloc organic64 fp32", W1)
        ids = np.arange(W1)[::-1].copy()
        ref = call(hip, c1, "device", ids, R1[ids])
        got = call(hip, c1, "noalloc", ids, R1[ids])
        check_equal(got, ref, "organic64 fp32")
        say("noalloc mol256 fp64 again", 5)
        ids = np.array([3, 11, 0, 20, 8])
        ref = call(hip, c2, "device", ids, R2[ids])
        got = call(hip, c2, "noalloc", ids, R2[ids])
        check_equal(got, ref, "mol256 fp64 after organic64")
