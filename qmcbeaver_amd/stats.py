"""Weighted DDDA statistics (binary blocking) over the native estimator.

`Property` mirrors the reference's QMCProperty (QMCProperty.cpp:127-237,
427-598, 830-872): `add` is newSample, `+`/`merge` is operator+, `reweight`
is reWeight, `average`, `std_dev`, `decorr_depth` and `correlation_length`
are the getters of the same meaning.  The arithmetic runs in
libqmcb_amd.so (host code, no device needed).

`series_energy(rows)` turns the per-sweep rows of Context.sample_series()
into the energy estimate the reference prints: one sample per sweep, the
walker-weighted average with the total weight (QMCRun.cpp:544-576).
"""
from __future__ import annotations

import ctypes as C
from typing import Optional

import numpy as np

from ._abi import PropertySummary
from .core import lib

LEVELS = 30


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


class Property:
    def __init__(self):
        self._h = lib().qmcb_property_create()
        if not self._h:
            raise MemoryError("qmcb_property_create")

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            lib().qmcb_property_destroy(h)
            self._h = None

    def add(self, sample: float, weight: float = 1.0) -> "Property":
        lib().qmcb_property_add(self._h, float(sample), float(weight))
        return self

    def add_series(self, samples, weights: Optional[np.ndarray] = None) -> "Property":
        s = np.ascontiguousarray(samples, dtype=np.float64).ravel()
        w = None
        if weights is not None:
            w = np.ascontiguousarray(weights, dtype=np.float64).ravel()
            if w.shape != s.shape:
                raise ValueError("samples and weights differ in length")
        st = lib().qmcb_property_add_series(self._h, s.size, _dp(s),
                                            None if w is None else _dp(w))
        if st != 0:
            raise ValueError("qmcb_property_add_series")
        return self

    def merge(self, other: "Property") -> "Property":
        if lib().qmcb_property_merge(self._h, other._h) != 0:
            raise ValueError("qmcb_property_merge")
        return self

    def __iadd__(self, other: "Property") -> "Property":
        return self.merge(other)

    def reweight(self, w: float) -> "Property":
        lib().qmcb_property_reweight(self._h, float(w))
        return self

    def zero(self) -> "Property":
        lib().qmcb_property_zero(self._h)
        return self

    def summary(self) -> PropertySummary:
        s = PropertySummary()
        if lib().qmcb_property_summary_get(self._h, C.byref(s)) != 0:
            raise ValueError("qmcb_property_summary_get")
        return s

    def level(self, i: int):
        """(samples, summed weights, average, block variance) of level i."""
        n = C.c_int64()
        w, a, v = C.c_double(), C.c_double(), C.c_double()
        if lib().qmcb_property_level(self._h, i, C.byref(n), C.byref(w), C.byref(a),
                                     C.byref(v)) != 0:
            raise IndexError(i)
        return n.value, w.value, a.value, v.value

    @property
    def n_samples(self) -> int:
        return self.summary().n_samples

    @property
    def average(self) -> float:
        return self.summary().average

    @property
    def std_dev(self) -> float:
        return self.summary().std_dev

    @property
    def decorr_depth(self) -> int:
        return self.summary().decorr_depth

    @property
    def correlation_length(self) -> float:
        return self.summary().correlation_length

    def __repr__(self):
        s = self.summary()
        return (f"Property(n={s.n_samples}, average={s.average:.10g}, std={s.std_dev:.3g}, "
                f"depth={s.decorr_depth})")


def series_energy(rows: np.ndarray, prop: Optional[Property] = None) -> Property:
    """Accumulate per-sweep rows (sum w E, sum w, ...) as the reference does:
    sample = sum w E / sum w, weight = sum w."""
    rows = np.asarray(rows, dtype=np.float64).reshape(-1, 4)
    ok = rows[:, 1] > 0
    p = prop if prop is not None else Property()
    p.add_series(rows[ok, 0] / rows[ok, 1], rows[ok, 1])
    return p
