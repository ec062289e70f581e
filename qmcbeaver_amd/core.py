"""Python host binding of the C ABI (include/qmcbeaver_amd.h) via ctypes.

The product path is the HIP library libqmcb_amd.so; this module only loads
it and marshals arrays.  There is no CPU fallback: if the library is missing
or cannot be loaded, `lib()` raises.
"""
from __future__ import annotations

import ctypes as C
import os
from typing import Dict, Optional, Sequence

import numpy as np

from ._abi import (Options, ResultArrays, Results, SystemBuffers, SystemStruct,
                   VmcStats, DmcParams, DmcStats, PropertySummary, CuspOrbital, PRECISIONS,
                   STATUS, QMCB_DRIFT_NONE)

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("QMCB_LIB_PATH") or os.path.join(HERE, "libqmcb_amd.so")

EXPORTS = ["qmcb_abi_version", "qmcb_create", "qmcb_destroy", "qmcb_last_error",
           "qmcb_evaluate", "qmcb_evaluate_device", "qmcb_evaluate_device_noalloc", "qmcb_accept",
           "qmcb_refresh",
           "qmcb_get_positions", "qmcb_vmc_steps", "qmcb_sync", "qmcb_kernel_times",
           "qmcb_enable_kernel_timing", "qmcb_dmc_init", "qmcb_dmc_steps",
           "qmcb_dmc_population", "qmcb_sample_series", "qmcb_property_create",
           "qmcb_property_destroy", "qmcb_property_zero", "qmcb_property_add",
           "qmcb_property_add_series", "qmcb_property_merge", "qmcb_property_reweight",
           "qmcb_property_summary_get", "qmcb_property_level", "qmcb_cusp_fit",
           "qmcb_device_bytes", "qmcb_plan_device_bytes", "qmcb_set_walker_groups",
           "qmcb_get_walker_groups"]

_lib = None


class QmcbError(RuntimeError):
    pass


def lib():
    """Load libqmcb_amd.so (raises if it was not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise QmcbError(f"{LIB_PATH} is missing: run __graft_entry__.build() "
                        "(hipcc --offload-arch=gfx950) first")
    L = C.CDLL(LIB_PATH)
    vp = C.c_void_p
    ip = C.POINTER(C.c_int32)
    dp = C.POINTER(C.c_double)
    L.qmcb_abi_version.restype = C.c_int32
    L.qmcb_create.restype = C.c_int
    L.qmcb_create.argtypes = [C.POINTER(SystemStruct), C.POINTER(Options), C.POINTER(vp)]
    L.qmcb_destroy.argtypes = [vp]
    L.qmcb_last_error.restype = C.c_char_p
    L.qmcb_last_error.argtypes = [vp]
    L.qmcb_evaluate.restype = C.c_int
    L.qmcb_evaluate.argtypes = [vp, C.c_int32, ip, dp, C.c_int32, C.POINTER(Results)]
    L.qmcb_evaluate_device.restype = C.c_int
    L.qmcb_evaluate_device.argtypes = [vp, C.c_int32, vp, vp, C.c_int32, C.POINTER(Results)]
    L.qmcb_evaluate_device_noalloc.restype = C.c_int
    L.qmcb_evaluate_device_noalloc.argtypes = [vp, C.c_int32, vp, vp, C.POINTER(Results)]
    L.qmcb_accept.restype = C.c_int
    L.qmcb_accept.argtypes = [vp, C.c_int32, ip, C.POINTER(C.c_uint8)]
    L.qmcb_refresh.restype = C.c_int
    L.qmcb_refresh.argtypes = [vp, C.c_int32, ip]
    L.qmcb_get_positions.restype = C.c_int
    L.qmcb_get_positions.argtypes = [vp, C.c_int32, dp]
    L.qmcb_vmc_steps.restype = C.c_int
    L.qmcb_vmc_steps.argtypes = [vp, C.c_int32, C.c_int64, C.c_int32, C.c_double,
                                 C.POINTER(VmcStats)]
    L.qmcb_sync.restype = C.c_int
    L.qmcb_sync.argtypes = [vp]
    L.qmcb_kernel_times.restype = C.c_int32
    L.qmcb_kernel_times.argtypes = [vp, C.c_int32, C.POINTER(C.c_char_p), dp,
                                    C.POINTER(C.c_int64)]
    L.qmcb_enable_kernel_timing.argtypes = [vp, C.c_int32]
    L.qmcb_dmc_init.restype = C.c_int
    L.qmcb_dmc_init.argtypes = [vp, C.c_int32]
    L.qmcb_dmc_steps.restype = C.c_int
    L.qmcb_dmc_steps.argtypes = [vp, C.c_int64, C.c_int32, C.POINTER(DmcParams),
                                 C.POINTER(DmcStats)]
    L.qmcb_dmc_population.restype = C.c_int
    L.qmcb_dmc_population.argtypes = [vp, C.c_int32, ip, dp, C.POINTER(C.c_int32)]
    L.qmcb_cusp_fit.restype = C.c_int
    L.qmcb_cusp_fit.argtypes = [C.POINTER(SystemStruct), C.c_int32, C.POINTER(CuspOrbital)]
    L.qmcb_set_walker_groups.restype = C.c_int
    L.qmcb_set_walker_groups.argtypes = [vp, C.c_int32]
    L.qmcb_get_walker_groups.restype = C.c_int
    L.qmcb_get_walker_groups.argtypes = [vp, C.POINTER(C.c_int32)]
    L.qmcb_device_bytes.restype = C.c_int
    L.qmcb_device_bytes.argtypes = [vp, C.POINTER(C.c_int64)]
    L.qmcb_plan_device_bytes.restype = C.c_int
    L.qmcb_plan_device_bytes.argtypes = [C.POINTER(SystemStruct), C.POINTER(Options),
                                         C.POINTER(C.c_int64)]
    L.qmcb_sample_series.restype = C.c_int
    L.qmcb_sample_series.argtypes = [vp, C.c_int32, dp, C.POINTER(C.c_int32)]
    L.qmcb_property_create.restype = vp
    L.qmcb_property_create.argtypes = []
    L.qmcb_property_destroy.argtypes = [vp]
    L.qmcb_property_zero.argtypes = [vp]
    L.qmcb_property_add.argtypes = [vp, C.c_double, C.c_double]
    L.qmcb_property_add_series.restype = C.c_int
    L.qmcb_property_add_series.argtypes = [vp, C.c_int64, dp, dp]
    L.qmcb_property_merge.restype = C.c_int
    L.qmcb_property_merge.argtypes = [vp, vp]
    L.qmcb_property_reweight.argtypes = [vp, C.c_double]
    L.qmcb_property_summary_get.restype = C.c_int
    L.qmcb_property_summary_get.argtypes = [vp, C.POINTER(PropertySummary)]
    L.qmcb_property_level.restype = C.c_int
    L.qmcb_property_level.argtypes = [vp, C.c_int32, C.POINTER(C.c_int64), dp, dp, dp]
    _lib = L
    return L


def cusp_fit(packed: Dict, spin: int = 0) -> np.ndarray:
    """Fitted cusp-replacement parameters of one spin's orbitals (host code,
    no device): structured array [n_nuc][n_orbitals] with fields rc, C,
    sgn0, alpha[5], phi0, n0, zeff, sigma_sq, ideal0."""
    buf = SystemBuffers(packed)
    nn = int(packed["n_nuc"])
    norb = int(packed["n_up"] if spin == 0 else packed["n_dn"])
    arr = (CuspOrbital * max(nn * norb, 1))()
    st = lib().qmcb_cusp_fit(C.byref(buf.struct), spin, arr)
    if st != 0:
        raise QmcbError(f"qmcb_cusp_fit: {STATUS.get(st, st)}")
    dt = np.dtype([("rc", "f8"), ("C", "f8"), ("sgn0", "f8"), ("alpha", "f8", 5), ("phi0", "f8"),
                   ("n0", "f8"), ("zeff", "f8"), ("sigma_sq", "f8"), ("ideal0", "f8")])
    out = np.frombuffer(bytes(arr), dtype=dt)[:nn * norb].copy()
    return out.reshape(nn, norb)


def _options(precision, max_walkers, drift_mod=QMCB_DRIFT_NONE, dt=0.1, umrigar_a=0.5,
             refresh_interval=0, seed=12345, device=0, replace_cusps=False,
             basis_scratch_bytes=0, walker_groups=0, old_walker_acceptance=50,
             all_electron_moves=False, graphs=False) -> Options:
    o = Options()
    o.precision = PRECISIONS[precision] if isinstance(precision, str) else int(precision)
    o.max_walkers = max_walkers
    o.drift_mod = drift_mod
    o.refresh_interval = refresh_interval
    o.dt = dt
    o.umrigar_a = umrigar_a
    o.seed = seed
    o.device = device
    o.replace_cusps = 1 if replace_cusps else 0
    o.basis_scratch_bytes = int(basis_scratch_bytes)
    o.walker_groups = int(walker_groups)
    o.old_walker_acceptance = int(old_walker_acceptance)
    o.all_electron_moves = 1 if all_electron_moves else 0
    o.graphs = 1 if graphs else 0
    return o


def plan_device_bytes(packed: Dict, *, precision: str = "fp64", max_walkers: int = 1,
                      replace_cusps: bool = False, basis_scratch_bytes: int = 0,
                      walker_groups: int = 0) -> int:
    """Device bytes a Context of this system and these options will hold
    after its first evaluate (host-only projection, no device needed)."""
    buf = SystemBuffers(packed)
    o = _options(precision, max_walkers, replace_cusps=replace_cusps,
                 basis_scratch_bytes=basis_scratch_bytes, walker_groups=walker_groups)
    b = C.c_int64()
    st = lib().qmcb_plan_device_bytes(C.byref(buf.struct), C.byref(o), C.byref(b))
    if st != 0:
        raise QmcbError(f"qmcb_plan_device_bytes: {STATUS.get(st, st)}")
    return int(b.value)


def _ip(a):
    return a.ctypes.data_as(C.POINTER(C.c_int32))


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


class Context:
    """One molecule + resident walkers on one GPU.

    Mirrors QMCFunctions (QMCFunctions.h:26-105): `evaluate(ids, R, whichE)`
    is QMCFunctions::evaluate(walkerData, xData, num) with the walker records
    resident on the device; `accept` commits single-electron proposals.
    """

    def __init__(self, packed: Dict, *, precision: str = "fp64", max_walkers: int = 1,
                 drift_mod: int = QMCB_DRIFT_NONE, dt: float = 0.1, umrigar_a: float = 0.5,
                 refresh_interval: int = 0, seed: int = 12345, device: int = 0,
                 replace_cusps: bool = False, basis_scratch_bytes: int = 0,
                 walker_groups: int = 0, old_walker_acceptance: int = 50,
                 all_electron_moves: bool = False, graphs: bool = False):
        """old_walker_acceptance: the reference's old_walker_acceptance_parameter
        (acceptance boost and DMC weight decay of walkers stuck for more than
        this many moves); 0 switches the rule off.  all_electron_moves: every
        electron moves at each vmc step (the reference's one_e_per_iter 0)."""
        self._buf = SystemBuffers(packed)
        self.N = self._buf.N
        self.n_up = int(packed["n_up"])
        self.max_walkers = max_walkers
        o = _options(precision, max_walkers, drift_mod, dt, umrigar_a, refresh_interval, seed,
                     device, replace_cusps, basis_scratch_bytes, walker_groups,
                     old_walker_acceptance, all_electron_moves, graphs)
        self._opt = o
        h = C.c_void_p()
        st = lib().qmcb_create(C.byref(self._buf.struct), C.byref(o), C.byref(h))
        if st != 0:
            raise QmcbError(f"qmcb_create: {STATUS.get(st, st)}: "
                            f"{lib().qmcb_last_error(None).decode()}")
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            lib().qmcb_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def _check(self, st, what):
        if st != 0:
            raise QmcbError(f"{what}: {STATUS.get(st, st)}: "
                            f"{lib().qmcb_last_error(self._h).decode()}")

    def evaluate(self, ids: Sequence[int], R: np.ndarray, whichE: int = -1) -> ResultArrays:
        ids = np.ascontiguousarray(ids, dtype=np.int32)
        R = np.ascontiguousarray(R, dtype=np.float64).reshape(len(ids), self.N, 3)
        res = ResultArrays(len(ids), self.N)
        self._check(lib().qmcb_evaluate(self._h, len(ids), _ip(ids), _dp(R), whichE,
                                        C.byref(res.struct)), "qmcb_evaluate")
        return res

    def accept(self, ids: Sequence[int], mask: Sequence[bool]):
        ids = np.ascontiguousarray(ids, dtype=np.int32)
        m = np.ascontiguousarray(mask, dtype=np.uint8)
        self._check(lib().qmcb_accept(self._h, len(ids), _ip(ids),
                                      m.ctypes.data_as(C.POINTER(C.c_uint8))), "qmcb_accept")

    def refresh(self, ids: Sequence[int]):
        ids = np.ascontiguousarray(ids, dtype=np.int32)
        self._check(lib().qmcb_refresh(self._h, len(ids), _ip(ids)), "qmcb_refresh")

    def positions(self, num: int) -> np.ndarray:
        R = np.zeros((num, self.N, 3))
        self._check(lib().qmcb_get_positions(self._h, num, _dp(R)), "qmcb_get_positions")
        return R

    def vmc_steps(self, num: int, first_step: int, n_steps: int, tau: float) -> VmcStats:
        st = VmcStats()
        self._check(lib().qmcb_vmc_steps(self._h, num, first_step, n_steps, tau,
                                         C.byref(st)), "qmcb_vmc_steps")
        return st

    def dmc_init(self, num: int):
        self._check(lib().qmcb_dmc_init(self._h, num), "qmcb_dmc_init")

    def dmc_steps(self, first_step: int, n_steps: int, tau: float, e_trial: float, *,
                  target: int = 0, pop_control: float = 1.0, branch_threshold: float = 2.0,
                  fusion_threshold: float = 0.45, reweight_after: int = 100,
                  energy_modification: bool = False, rel_cutoff: float = 100.0,
                  equilibration_steps: int = 0) -> DmcStats:
        """energy_modification: the reference's umrigar93 energy modification
        of the weights (meaningful with drift_mod=QMCB_DRIFT_UMRIGAR93_UNEQUAL).
        rel_cutoff: delete walkers whose energy is further than rel_cutoff x
        |E_0| from E_0 (the first e_trial; 0: off).  equilibration_steps: the
        reference's equilibration phase (its trial-energy update and
        stricter deletion rules)."""
        p = DmcParams(tau, e_trial, pop_control, branch_threshold, fusion_threshold, target,
                      reweight_after, 1 if energy_modification else 0, rel_cutoff,
                      equilibration_steps)
        st = DmcStats()
        self._check(lib().qmcb_dmc_steps(self._h, first_step, n_steps, C.byref(p),
                                         C.byref(st)), "qmcb_dmc_steps")
        return st

    def sample_series(self) -> np.ndarray:
        """Per-sweep observable rows of the last vmc_steps/dmc_steps call:
        [rows, 4] = (sum w E, sum w, sum w p, walkers); the sweep's sample is
        row[0] / row[1] with weight row[1]."""
        n = C.c_int32()
        self._check(lib().qmcb_sample_series(self._h, 0, None, C.byref(n)),
                    "qmcb_sample_series")
        rows = np.zeros((max(n.value, 1), 4))
        self._check(lib().qmcb_sample_series(self._h, n.value, _dp(rows), C.byref(n)),
                    "qmcb_sample_series")
        return rows[:n.value].copy()

    def dmc_population(self):
        cap = self.max_walkers
        slots = np.zeros(cap, dtype=np.int32)
        w = np.zeros(cap)
        n = C.c_int32()
        self._check(lib().qmcb_dmc_population(self._h, cap, _ip(slots), _dp(w), C.byref(n)),
                    "qmcb_dmc_population")
        return slots[:n.value].copy(), w[:n.value].copy()

    def sync(self):
        self._check(lib().qmcb_sync(self._h), "qmcb_sync")

    def set_walker_groups(self, groups: int):
        """Walker groups of the next vmc_steps calls (results do not change);
        -1: chosen by timing at the next call with enough steps."""
        self._check(lib().qmcb_set_walker_groups(self._h, int(groups)), "qmcb_set_walker_groups")

    def walker_groups(self) -> int:
        """Walker-group count in use (with -1: the timed choice, 0 before one)."""
        g = C.c_int32(0)
        self._check(lib().qmcb_get_walker_groups(self._h, C.byref(g)), "qmcb_get_walker_groups")
        return int(g.value)

    def device_bytes(self) -> int:
        """Device memory this context holds."""
        b = C.c_int64()
        self._check(lib().qmcb_device_bytes(self._h, C.byref(b)), "qmcb_device_bytes")
        return int(b.value)

    def enable_timing(self, on: bool = True):
        lib().qmcb_enable_kernel_timing(self._h, 1 if on else 0)

    def kernel_times(self) -> Dict[str, Dict[str, float]]:
        cap = 32
        names = (C.c_char_p * cap)()
        ms = np.zeros(cap)
        n = np.zeros(cap, dtype=np.int64)
        k = lib().qmcb_kernel_times(self._h, cap, names, _dp(ms),
                                    n.ctypes.data_as(C.POINTER(C.c_int64)))
        return {names[i].decode(): {"ms": float(ms[i]), "launches": int(n[i])}
                for i in range(min(k, cap))}
