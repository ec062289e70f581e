"""Flat binary system/walker files read by the C++ drop-in example."""
from __future__ import annotations

import numpy as np


def write_flat(packed, path: str) -> None:
    with open(path, "wb") as f:
        np.array([packed["n_up"], packed["n_dn"], packed["n_nuc"], packed["n_basis"],
                  packed["n_prim"]], dtype=np.int32).tofile(f)
        for k, dt in (("nuc_xyz", np.float64), ("nuc_z", np.float64), ("bf_atom", np.int32),
                      ("bf_pow", np.int32), ("bf_prim_off", np.int32), ("prim_alpha", np.float64),
                      ("prim_coef", np.float64), ("mo_up", np.float64), ("mo_dn", np.float64)):
            np.ascontiguousarray(packed[k], dtype=dt).tofile(f)
        rows = [packed["ee_updn"], packed["ee_upup"], packed["ee_dndn"],
                *packed["en_up"], *packed["en_dn"]]
        for r in rows:  # qmcb_pade: int32 n_num, n_den, double num[8], den[8]
            np.array([int(r[0]), int(r[1])], dtype=np.int32).tofile(f)
            np.asarray(r[2:18], dtype=np.float64).tofile(f)


def write_walkers(R: np.ndarray, path: str) -> None:
    with open(path, "wb") as f:
        np.array(R.shape[:2], dtype=np.int32).tofile(f)
        np.ascontiguousarray(R, dtype=np.float64).tofile(f)
