"""ctypes front end of the C oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, never by the product package.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

from qmcbeaver_amd._abi import ResultArrays, SystemBuffers, Results, SystemStruct, Pade

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
_lib = None


def build() -> None:
    subprocess.run(["make", "-C", HERE, "liboracle.so"], check=True,
                   stdout=subprocess.DEVNULL)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        L.orc_walker_new.restype = C.c_void_p
        L.orc_walker_new.argtypes = [C.POINTER(SystemStruct)]
        L.orc_walker_free.argtypes = [C.c_void_p]
        L.orc_walker_copy.argtypes = [C.c_void_p, C.c_void_p]
        L.orc_evaluate.restype = C.c_int
        L.orc_evaluate.argtypes = [C.POINTER(SystemStruct), C.c_void_p,
                                   C.POINTER(C.c_double), C.c_int, C.c_int,
                                   C.c_double, C.c_double, C.POINTER(Results)]
        dp = C.POINTER(C.c_double)
        L.orc_basis.argtypes = [C.POINTER(SystemStruct), dp, dp, dp, dp, dp, dp]
        L.orc_pade.argtypes = [C.POINTER(Pade), C.c_double, dp, dp, dp]
        L.orc_det_inverse.restype = C.c_int
        L.orc_det_inverse.argtypes = [C.c_int, dp, dp, dp]
        L.orc_inverse_update_row.restype = C.c_double
        L.orc_inverse_update_row.argtypes = [C.c_int, dp, C.c_int, dp]
        L.orc_vmc_time.restype = C.c_double
        L.orc_vmc_time.argtypes = [C.POINTER(SystemStruct), C.c_int, dp, C.c_int,
                                   C.c_double, C.c_ulonglong, dp]
        _lib = L
    return _lib


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


class Oracle:
    """One system; walkers are oracle records (QMCWalkerData analogues)."""

    def __init__(self, packed, drift_mod=0, dt=0.1, umrigar_a=0.5):
        self.buf = SystemBuffers(packed)
        self.N = self.buf.N
        self.drift_mod, self.dt, self.umrigar_a = drift_mod, dt, umrigar_a
        self.walkers = []

    def __del__(self):
        try:
            for w in self.walkers:
                lib().orc_walker_free(w)
        except Exception:
            pass

    def new_walker(self):
        w = lib().orc_walker_new(C.byref(self.buf.struct))
        self.walkers.append(w)
        return w

    def evaluate(self, w, R, whichE=-1, res: ResultArrays | None = None, row=0):
        R = np.ascontiguousarray(R, dtype=np.float64).reshape(self.N, 3)
        if res is None:
            res = ResultArrays(1, self.N)
        r = res.row(row)
        lib().orc_evaluate(C.byref(self.buf.struct), w, _dp(R), whichE,
                           self.drift_mod, self.dt, self.umrigar_a, C.byref(r))
        return res

    def evaluate_all(self, R):
        """Fresh walker per configuration, all-electron evaluate."""
        R = np.asarray(R, dtype=np.float64)
        res = ResultArrays(R.shape[0], self.N)
        for i in range(R.shape[0]):
            w = lib().orc_walker_new(C.byref(self.buf.struct))
            Ri = np.ascontiguousarray(R[i])
            r = res.row(i)
            lib().orc_evaluate(C.byref(self.buf.struct), w, _dp(Ri), -1,
                               self.drift_mod, self.dt, self.umrigar_a, C.byref(r))
            lib().orc_walker_free(w)
        return res

    def basis(self, r):
        B = int(self.buf.struct.n_basis)
        out = [np.zeros(B) for _ in range(5)]
        r = np.ascontiguousarray(r, dtype=np.float64)
        lib().orc_basis(C.byref(self.buf.struct), _dp(r), *[_dp(o) for o in out])
        return out

    def vmc_time(self, R0, n_moves, tau, seed=1):
        R0 = np.ascontiguousarray(R0, dtype=np.float64)
        mean = C.c_double(0.0)
        t = lib().orc_vmc_time(C.byref(self.buf.struct), R0.shape[0], _dp(R0),
                               n_moves, tau, seed, C.byref(mean))
        return t, mean.value


def pade(row, r):
    p = Pade()
    p.n_num, p.n_den = int(row[0]), int(row[1])
    for i in range(8):
        p.num[i] = float(row[2 + i]); p.den[i] = float(row[10 + i])
    u, du, d2u = C.c_double(), C.c_double(), C.c_double()
    lib().orc_pade(C.byref(p), float(r), C.byref(u), C.byref(du), C.byref(d2u))
    return u.value, du.value, d2u.value


def det_inverse(A):
    A = np.array(A, dtype=np.float64, order="C")
    n = A.shape[0]
    inv = np.zeros_like(A)
    det = C.c_double()
    ok = lib().orc_det_inverse(n, _dp(A), _dp(inv), C.byref(det))
    return ok, inv, det.value


def inverse_update_row(inv, row, newrow):
    inv = np.array(inv, dtype=np.float64, order="C")
    nr = np.ascontiguousarray(newrow, dtype=np.float64)
    ratio = lib().orc_inverse_update_row(inv.shape[0], _dp(inv), row, _dp(nr))
    return ratio, inv
