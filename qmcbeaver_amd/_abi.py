"""ctypes mirror of include/qmcbeaver_amd.h (structs only, no library load)."""
from __future__ import annotations

import ctypes as C
from typing import Dict

import numpy as np

PADE_MAX = 8

QMCB_FP64, QMCB_MIXED_FP32, QMCB_MIXED_BF16, QMCB_MIXED_F32STORE, QMCB_MIXED_BF16STORE = \
    0, 1, 2, 3, 4
QMCB_DRIFT_NONE, QMCB_DRIFT_UMRIGAR93_UNEQUAL, QMCB_DRIFT_UMRIGAR93_EQUAL = 0, 1, 2
PRECISIONS = {"fp64": QMCB_FP64, "fp32": QMCB_MIXED_FP32, "bf16": QMCB_MIXED_BF16,
              "fp32store": QMCB_MIXED_F32STORE, "bf16store": QMCB_MIXED_BF16STORE}
STATUS = {0: "OK", 1: "INVALID_ARG", 2: "OOM", 3: "DEVICE_ERROR", 4: "BAD_STATE"}

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int32)
_i8p = C.POINTER(C.c_int8)


class Pade(C.Structure):
    _fields_ = [("n_num", C.c_int32), ("n_den", C.c_int32),
                ("num", C.c_double * PADE_MAX), ("den", C.c_double * PADE_MAX)]


class SystemStruct(C.Structure):
    _fields_ = [
        ("n_up", C.c_int32), ("n_dn", C.c_int32), ("n_nuc", C.c_int32),
        ("n_basis", C.c_int32), ("n_prim", C.c_int32),
        ("nuc_xyz", _dp), ("nuc_z", _dp),
        ("bf_atom", _ip), ("bf_pow", _ip), ("bf_prim_off", _ip),
        ("prim_alpha", _dp), ("prim_coef", _dp),
        ("mo_up", _dp), ("mo_dn", _dp),
        ("ee_updn", Pade), ("ee_upup", Pade), ("ee_dndn", Pade),
        ("en_up", C.POINTER(Pade)), ("en_dn", C.POINTER(Pade)),
        ("n_det", C.c_int32), ("n_act_up", C.c_int32), ("n_act_dn", C.c_int32),
        ("ci_coef", _dp), ("occ_up", _ip), ("occ_dn", _ip),
    ]


class Options(C.Structure):
    _fields_ = [("precision", C.c_int32), ("max_walkers", C.c_int32),
                ("drift_mod", C.c_int32), ("refresh_interval", C.c_int32),
                ("dt", C.c_double), ("umrigar_a", C.c_double),
                ("seed", C.c_uint64), ("device", C.c_int32), ("replace_cusps", C.c_int32),
                ("basis_scratch_bytes", C.c_int64), ("walker_groups", C.c_int32),
                ("old_walker_acceptance", C.c_int32), ("all_electron_moves", C.c_int32),
                ("graphs", C.c_int32)]


class Results(C.Structure):
    _fields_ = [("log_psi", _dp), ("sign_psi", _i8p), ("grad_lnpsi", _dp),
                ("mod_drift", _dp), ("lap_ratio", _dp), ("e_local", _dp),
                ("kinetic", _dp), ("potential", _dp), ("v_ne", _dp), ("v_ee", _dp),
                ("det_ratio", _dp), ("jastrow_u", _dp), ("singular", _i8p),
                ("mod_ratio", _dp), ("r2", _dp)]


class CuspOrbital(C.Structure):
    _fields_ = [("rc", C.c_double), ("C", C.c_double), ("sgn0", C.c_double),
                ("alpha", C.c_double * 5), ("phi0", C.c_double), ("n0", C.c_double),
                ("zeff", C.c_double), ("sigma_sq", C.c_double), ("ideal0", C.c_double)]


class VmcStats(C.Structure):
    _fields_ = [("sum_e_local", C.c_double), ("sum_e_local_sq", C.c_double),
                ("sum_accept", C.c_double), ("sum_prop_e_local", C.c_double),
                ("n_samples", C.c_int64)]


class DmcParams(C.Structure):
    _fields_ = [("tau", C.c_double), ("e_trial", C.c_double), ("pop_control", C.c_double),
                ("branch_threshold", C.c_double), ("fusion_threshold", C.c_double),
                ("target_walkers", C.c_int32), ("reweight_after", C.c_int32),
                ("energy_modification", C.c_int32), ("rel_cutoff", C.c_double),
                ("equilibration_steps", C.c_int32)]


class DmcStats(C.Structure):
    _fields_ = [("sum_w_e_local", C.c_double), ("sum_w_e_local_sq", C.c_double),
                ("sum_weights", C.c_double), ("sum_accept", C.c_double),
                ("n_moves", C.c_int64), ("population", C.c_int32), ("n_split", C.c_int32),
                ("n_killed", C.c_int32), ("n_fused", C.c_int32), ("total_weight", C.c_double),
                ("e_trial", C.c_double), ("n_split_skipped", C.c_int32),
                ("n_deleted", C.c_int32), ("e_estimated", C.c_double)]


class PropertySummary(C.Structure):
    _fields_ = [("n_samples", C.c_int64), ("average", C.c_double), ("std_dev", C.c_double),
                ("std_dev_std", C.c_double), ("correlation_length", C.c_double),
                ("serial_variance", C.c_double), ("skewness", C.c_double),
                ("kurtosis", C.c_double), ("decorr_depth", C.c_int32), ("n_levels", C.c_int32)]


def _pade_struct(row) -> Pade:
    p = Pade()
    p.n_num, p.n_den = int(row[0]), int(row[1])
    for i in range(PADE_MAX):
        p.num[i] = float(row[2 + i])
        p.den[i] = float(row[2 + PADE_MAX + i])
    return p


class SystemBuffers:
    """Keeps the packed numpy arrays alive for the lifetime of the struct."""

    def __init__(self, packed: Dict[str, np.ndarray]):
        self.arrays = {}
        for k, v in packed.items():
            if isinstance(v, np.ndarray):
                self.arrays[k] = np.ascontiguousarray(v)
        a = self.arrays
        s = SystemStruct()
        for k in ("n_up", "n_dn", "n_nuc", "n_basis", "n_prim"):
            setattr(s, k, int(packed[k]))
        for k in ("nuc_xyz", "nuc_z", "prim_alpha", "prim_coef", "mo_up", "mo_dn"):
            assert a[k].dtype == np.float64
            setattr(s, k, a[k].ctypes.data_as(_dp))
        for k in ("bf_atom", "bf_pow", "bf_prim_off"):
            assert a[k].dtype == np.int32
            setattr(s, k, a[k].ctypes.data_as(_ip))
        s.ee_updn = _pade_struct(a["ee_updn"])
        s.ee_upup = _pade_struct(a["ee_upup"])
        s.ee_dndn = _pade_struct(a["ee_dndn"])
        n = int(packed["n_nuc"])
        self.en_up = (Pade * n)(*[_pade_struct(r) for r in a["en_up"]])
        self.en_dn = (Pade * n)(*[_pade_struct(r) for r in a["en_dn"]])
        s.en_up = C.cast(self.en_up, C.POINTER(Pade))
        s.en_dn = C.cast(self.en_dn, C.POINTER(Pade))
        if int(packed.get("n_det", 1)) > 1 or "occ_up" in packed:
            s.n_det = int(packed["n_det"])
            s.n_act_up = int(packed["n_act_up"])
            s.n_act_dn = int(packed["n_act_dn"])
            for k in ("ci_coef",):
                assert a[k].dtype == np.float64
                setattr(s, k, a[k].ctypes.data_as(_dp))
            for k in ("occ_up", "occ_dn"):
                assert a[k].dtype == np.int32
                setattr(s, k, a[k].ctypes.data_as(_ip))
        else:
            s.n_det = 1
        self.struct = s
        self.N = int(packed["n_up"]) + int(packed["n_dn"])


class ResultArrays:
    """Host numpy output arrays plus the Results struct pointing at them."""

    FIELDS = ("log_psi", "sign_psi", "grad_lnpsi", "mod_drift", "lap_ratio",
              "e_local", "kinetic", "potential", "v_ne", "v_ee", "det_ratio",
              "jastrow_u", "singular", "mod_ratio", "r2")

    def __init__(self, num: int, N: int):
        self.num, self.N = num, N
        self.a = {}
        for f in self.FIELDS:
            if f in ("sign_psi", "singular"):
                self.a[f] = np.zeros(num, dtype=np.int8)
            elif f in ("grad_lnpsi", "mod_drift"):
                self.a[f] = np.zeros((num, N, 3))
            elif f == "r2":
                self.a[f] = np.zeros((num, 3))
            else:
                self.a[f] = np.zeros(num)
        r = Results()
        for f in self.FIELDS:
            arr = self.a[f]
            setattr(r, f, arr.ctypes.data_as(_i8p if arr.dtype == np.int8 else _dp))
        self.struct = r

    def row(self, i: int) -> "Results":
        """Results struct addressing row i only (for single-walker oracle calls)."""
        r = Results()
        for f in self.FIELDS:
            arr = self.a[f]
            ptr = arr[i:i + 1] if arr.ndim == 1 else arr[i]
            setattr(r, f, ptr.ctypes.data_as(_i8p if arr.dtype == np.int8 else _dp))
        return r

    def __getitem__(self, k):
        return self.a[k]
