"""Seeded synthetic molecules for the walker-evaluation path.

A `System` holds exactly what the reference reads from a `.ckmf` input file:
nuclei (label, Z, position), per-atom Cartesian contracted Gaussians with raw
(un-normalised) contraction coefficients, the MO coefficient matrix with a
single restricted determinant occupying orbitals 0..N_sigma-1 for each spin,
and Pade two-body Jastrow functions u(r) = (p1 r + p2 r^2 ...)/(1 + q1 r + ...).

Conventions followed (reference file:line):
  * basis functions ordered by atom (geometry order), then file order, with
    contraction coefficients used as given (QMCBasisFunction.cpp:177-201,
    QMCBasisFunction.cpp:287-403);
  * Cartesian type strings -> powers (QMCBasisFunctionCoefficients.cpp:39-80);
  * per-spin coefficient matrices are the occupied rows of the MO matrix in
    orbital order (QMCWavefunction.cpp:873-905);
  * Pade parameter split: numerator p_1.. first, denominator q_1.. second
    (PadeCorrelationFunction.cpp:15-51);
  * electron order alpha first, then beta (QMCSCFJastrow.cpp:61-76).

`write_ckmf` emits the same system as a reference input file so that the
reference build and this project read byte-identical parameters.
"""
from __future__ import annotations

import dataclasses
import hashlib
import math
from typing import Dict, List, Tuple

import numpy as np

LABEL_OF_Z = {1: "H", 6: "C", 7: "N", 8: "O"}
CART_TYPES = {
    0: ["s"],
    1: ["px", "py", "pz"],
    2: ["dxx", "dyy", "dzz", "dxy", "dxz", "dyz"],
}
PADE_MAX = 8


def type_powers(t: str) -> Tuple[int, int, int]:
    """Cartesian powers (a, b, c) of a type string such as 'dxy'."""
    return (t[1:].count("x"), t[1:].count("y"), t[1:].count("z"))


@dataclasses.dataclass
class Shell:
    l: int
    alpha: np.ndarray  # primitive exponents
    coef: np.ndarray   # raw contraction coefficients


@dataclasses.dataclass
class System:
    name: str
    labels: List[str]
    Z: np.ndarray            # (Nn,) int
    xyz: np.ndarray          # (Nn, 3) bohr
    shells: List[List[Shell]]  # per atom
    mo: np.ndarray           # (norb, B) row = orbital
    n_up: int
    n_dn: int
    charge: int
    # Pade functions as (numerator p_1.., denominator q_1..)
    ee_updn: Tuple[List[float], List[float]]
    ee_upup: Tuple[List[float], List[float]]
    en: Dict[str, Tuple[List[float], List[float]]]  # per nucleus label

    # ------------------------------------------------------------------
    @property
    def n_elec(self) -> int:
        return self.n_up + self.n_dn

    @property
    def n_nuc(self) -> int:
        return len(self.labels)

    def basis_functions(self):
        """Yield (atom, type_string, alpha, coef) in reference BF order."""
        for a, shells in enumerate(self.shells):
            for sh in shells:
                for t in CART_TYPES[sh.l]:
                    yield a, t, sh.alpha, sh.coef

    @property
    def n_basis(self) -> int:
        return sum(len(CART_TYPES[sh.l]) for s in self.shells for sh in s)

    def pack(self) -> Dict[str, np.ndarray]:
        """Flat arrays in the layout of `qmcb_system` (include/qmcbeaver_amd.h)."""
        bf_atom, bf_pow, off, alpha, coef = [], [], [0], [], []
        for a, t, al, co in self.basis_functions():
            bf_atom.append(a)
            bf_pow.append(type_powers(t))
            alpha.extend(al.tolist())
            coef.extend(co.tolist())
            off.append(len(alpha))
        pade = lambda pq: _pade_arr(pq)
        en_up = np.stack([pade(self.en[l]) for l in self.labels])
        return {
            "n_up": self.n_up, "n_dn": self.n_dn, "n_nuc": self.n_nuc,
            "n_basis": len(bf_atom), "n_prim": len(alpha),
            "nuc_xyz": np.ascontiguousarray(self.xyz, dtype=np.float64),
            "nuc_z": np.asarray(self.Z, dtype=np.float64),
            "bf_atom": np.asarray(bf_atom, dtype=np.int32),
            "bf_pow": np.asarray(bf_pow, dtype=np.int32).reshape(-1, 3),
            "bf_prim_off": np.asarray(off, dtype=np.int32),
            "prim_alpha": np.asarray(alpha, dtype=np.float64),
            "prim_coef": np.asarray(coef, dtype=np.float64),
            "mo_up": np.ascontiguousarray(self.mo[: self.n_up], dtype=np.float64),
            "mo_dn": np.ascontiguousarray(self.mo[: self.n_dn], dtype=np.float64),
            "ee_updn": pade(self.ee_updn),
            "ee_upup": pade(self.ee_upup),
            # link_Jastrow_parameters 1: down-down = up-up, down-nucleus = up-nucleus
            "ee_dndn": pade(self.ee_upup),
            "en_up": en_up,
            "en_dn": en_up.copy(),
        }

    def digest(self) -> str:
        """SHA-256 of the packed arrays; golden fixtures record it."""
        h = hashlib.sha256()
        for k, v in sorted(self.pack().items()):
            h.update(k.encode())
            h.update(np.ascontiguousarray(np.asarray(v)).tobytes())
        return h.hexdigest()


def _pade_arr(pq) -> np.ndarray:
    """[n_num, n_den, num[PADE_MAX], den[PADE_MAX]] as float64 row."""
    p, q = pq
    assert len(p) <= PADE_MAX and len(q) <= PADE_MAX
    out = np.zeros(2 + 2 * PADE_MAX)
    out[0], out[1] = len(p), len(q)
    out[2:2 + len(p)] = p
    out[2 + PADE_MAX:2 + PADE_MAX + len(q)] = q
    return out


# ----------------------------------------------------------------------------
# generators
# ----------------------------------------------------------------------------

def _loguniform(rng, lo, hi, n):
    return np.exp(rng.uniform(math.log(lo), math.log(hi), size=n))


def _chain(rng, n, lo, hi):
    """Random chain: consecutive spacing U[lo,hi], every pair >= lo."""
    pts = [np.zeros(3)]
    while len(pts) < n:
        for _ in range(10000):
            d = rng.normal(size=3)
            d /= np.linalg.norm(d)
            p = pts[-1] + d * rng.uniform(lo, hi)
            if all(np.linalg.norm(p - q) >= lo for q in pts[:-1]):
                pts.append(p)
                break
        else:  # pragma: no cover - astronomically unlikely
            raise RuntimeError("chain generation failed")
    xyz = np.array(pts)
    return xyz - xyz.mean(axis=0)


def _orthonormal_rows(rng, nrow, ncol):
    q, _ = np.linalg.qr(rng.normal(size=(ncol, nrow)))
    return np.ascontiguousarray(q.T)


def _random_shell(rng, l, nprim, alo, ahi):
    return Shell(l, _loguniform(rng, alo, ahi, nprim), rng.uniform(0.1, 1.0, nprim))


def _check_exponent_boundaries(shells):
    """The reference reuses exp(-a r^2) across an atom boundary when the first
    exponent of atom k+1 equals the last one of atom k (QMCBasisFunction.cpp:
    310,372-379); such inputs must not be generated."""
    for a in range(len(shells) - 1):
        last = shells[a][-1].alpha[-1]
        first = shells[a + 1][0].alpha[0]
        assert abs(last - first) > 1e-8, "exponent shared across atom boundary"


PADE_DEFAULT = dict(
    ee_updn=([0.5], [1.0]), ee_upup=([0.25], [1.0]), en=([-1.0], [1.0]))


def make_system(config: int, seed: int | None = None, **kw) -> System:
    """Synthetic systems of the workload shapes:

    0: H2-like, 2 H at (0,0,+-0.7), 1+1 electrons, per atom 3 s + 1 p shell
       of 3 primitives, alpha ~ logU[0.1,50] -> 12 BFs; one occupied orbital
       (a random unit row of length 12).  norbitals always equals the number
       of occupied orbitals: with more orbitals than electrons the reference's
       one_e_per_iter row copy (Array2D::setRows, Array2D.h:964-987, used at
       QMCSlater.cpp:486-490) mixes strides.
    1: 4 C + 2 N + 4 H chain (spacing U[1.4,2.0]), 32+32 electrons, charge
       -22 as stated, heavy 7s5p3d (40 BF) / H 6s3p (15 BF) -> 300 BFs,
       3-6 primitives, alpha ~ logU[0.05,1e4].
    2: 4 H + 12 C + 12 N + 12 O chain (U[1.4,2.2]), neutral, 128+128,
       heavy 5s5p2d (32 BF) / H 3s3p (12 BF) -> 1200 BFs.
    'mol' (config=-1): keyword-driven molecule for batched/serving workloads:
       n_elec even, B ~ 5 n_elec.
    """
    seed = (0x51EED00 + max(config, 0)) if seed is None else seed
    rng = np.random.default_rng(seed)
    pade = dict(PADE_DEFAULT)
    if config == 0:
        Z = np.array([1, 1])
        xyz = np.array([[0.0, 0.0, 0.7], [0.0, 0.0, -0.7]])
        shells = [[_random_shell(rng, 0, 3, 0.1, 50.0) for _ in range(3)]
                  + [_random_shell(rng, 1, 3, 0.1, 50.0)] for _ in range(2)]
        n_up = n_dn = 1
        norb = 1
        name = "h2_synth"
    elif config in (1, 2):
        if config == 1:
            Z = np.array([6] * 4 + [7] * 2 + [1] * 4)
            heavy, hyd, lo, hi = (7, 5, 3), (6, 3, 0), 1.4, 2.0
            n_up = n_dn = 32
            name = "organic64_synth"
        else:
            Z = np.array([1] * 4 + [6] * 12 + [7] * 12 + [8] * 12)
            heavy, hyd, lo, hi = (5, 5, 2), (3, 3, 0), 1.4, 2.2
            n_up = n_dn = 128
            name = "mol256_synth"
        Z = rng.permutation(Z)
        xyz = _chain(rng, len(Z), lo, hi)
        shells = []
        for z in Z:
            ns = heavy if z > 1 else hyd
            sh = []
            for l, cnt in enumerate(ns):
                for _ in range(cnt):
                    sh.append(_random_shell(rng, l, int(rng.integers(3, 7)), 0.05, 1e4))
            shells.append(sh)
        norb = n_up
    elif config == -1:
        n_elec = int(kw["n_elec"])
        assert n_elec % 2 == 0 and n_elec >= 2
        Zs = []
        while sum(Zs) < n_elec:
            Zs.append(int(rng.choice([1, 6, 7, 8])))
        Zs[-1] -= sum(Zs) - n_elec
        if Zs[-1] not in LABEL_OF_Z:
            Zs[-1] = 1 if Zs[-1] < 6 else 6
            while sum(Zs) < n_elec:
                Zs.append(1)
        Z = np.array(Zs)
        xyz = _chain(rng, len(Z), 1.4, 2.2)
        target_b = int(kw.get("basis_per_elec", 5) * n_elec)
        per_atom = max(1, target_b // len(Z))
        shells = []
        for z in Z:
            sh, cnt = [], 0
            while cnt < per_atom:
                l = int(rng.choice([0, 1, 2], p=[0.4, 0.4, 0.2]))
                sh.append(_random_shell(rng, l, int(rng.integers(3, 7)), 0.05, 1e4))
                cnt += len(CART_TYPES[l])
            sh.sort(key=lambda s: s.l)
            shells.append(sh)
        n_up = n_dn = n_elec // 2
        norb = n_up
        name = f"mol{n_elec}_synth"
    else:
        raise ValueError(config)
    _check_exponent_boundaries(shells)
    labels = [LABEL_OF_Z[int(z)] for z in Z]
    B = sum(len(CART_TYPES[sh.l]) for s in shells for sh in s)
    mo = _orthonormal_rows(rng, norb, B)
    charge = int(Z.sum()) - (n_up + n_dn)
    en = {l: pade["en"] for l in set(labels)}
    return System(name=name, labels=labels, Z=Z.astype(np.int64), xyz=xyz,
                  shells=shells, mo=mo, n_up=n_up, n_dn=n_dn, charge=charge,
                  ee_updn=pade["ee_updn"], ee_upup=pade["ee_upup"], en=en)


def initial_walkers(sys_: System, n_walkers: int, seed: int, sigma: float = 0.5) -> np.ndarray:
    """R[w][e][3] ~ N(R_A, sigma^2 I); electrons assigned to nuclei in
    proportion to Z, alpha and beta electrons interleaved over that list."""
    rng = np.random.default_rng(seed)
    order = np.repeat(np.arange(sys_.n_nuc), np.maximum(sys_.Z, 1))
    N = sys_.n_elec
    nuc = np.empty(N, dtype=np.int64)
    for k in range(sys_.n_up):
        nuc[k] = order[(2 * k) % len(order)]
    for k in range(sys_.n_dn):
        nuc[sys_.n_up + k] = order[(2 * k + 1) % len(order)]
    R = sys_.xyz[nuc][None, :, :] + sigma * rng.normal(size=(n_walkers, N, 3))
    return np.ascontiguousarray(R)


# ----------------------------------------------------------------------------
# reference input file
# ----------------------------------------------------------------------------

def _fmt(x: float) -> str:
    return repr(float(x))


def write_ckmf(sys_: System, path: str, *, walkers: int = 1, dt: float = 0.1,
               one_e_per_iter: int = 1, energy: float = -1.0,
               run_type: str = "variational", steps: int = 10,
               extra_flags: Dict[str, str] | None = None) -> None:
    """Write `sys_` as a reference `.ckmf` input (flags for a plain-Pade,
    single-determinant, no-cusp-replacement, importance-sampled run)."""
    B = sys_.n_basis
    norb = sys_.mo.shape[0]
    flags = {
        "atoms": str(sys_.n_nuc), "charge": str(sys_.charge),
        "energy": _fmt(energy), "norbitals": str(norb), "nbasisfunc": str(B),
        "ndeterminants": "1", "trial_function_type": "restricted",
        "run_type": run_type, "iseed": "-123456", "dt": _fmt(dt),
        "dt_equilibration": _fmt(dt), "max_time_steps": str(steps),
        "equilibration_steps": "0", "number_of_walkers": str(walkers),
        "walkers_per_pass": "1", "one_e_per_iter": str(one_e_per_iter),
        "replace_electron_nucleus_cusps": "0", "use_three_body_jastrow": "0",
        "sampling_method": "importance_sampling", "QF_modification_type": "none",
        "energy_cutoff_type": "none", "energy_modification_type": "none",
        "output_interval": "1", "use_available_checkpoints": "0",
        "link_Jastrow_parameters": "1", "chip_and_mike_are_cool": "Yea_Baby!",
    }
    if extra_flags:
        flags.update(extra_flags)
    L = ["&flags"]
    for k, v in flags.items():
        L += [k, " " + v]
    L += ["&", "&geometry"]
    for lab, z, p in zip(sys_.labels, sys_.Z, sys_.xyz):
        L.append(f"{lab}\t{int(z)}\t{_fmt(p[0])}\t{_fmt(p[1])}\t{_fmt(p[2])}")
    L += ["&", "&basis"]
    for a, shells in enumerate(sys_.shells):
        nbf = sum(len(CART_TYPES[sh.l]) for sh in shells)
        maxg = max(len(sh.alpha) for sh in shells)
        L.append(f"{sys_.labels[a]}\t{nbf}\t{maxg}")
        for sh in shells:
            for t in CART_TYPES[sh.l]:
                L.append(f"\t{len(sh.alpha)}\t{t}")
                for al, co in zip(sh.alpha, sh.coef):
                    L.append(f"\t\t{_fmt(al)}\t{_fmt(co)}")
    L += ["&", "&wavefunction"]
    for row in sys_.mo:
        L.append("\t".join(_fmt(x) for x in row))
        L.append("")
    occ_a = ["1" if i < sys_.n_up else "0" for i in range(norb)]
    occ_b = ["1" if i < sys_.n_dn else "0" for i in range(norb)]
    L += ["Alpha Occupation", "\t".join(occ_a), "", "Beta Occupation",
          "\t".join(occ_b), "", "CI Coeffs", "1", "", "&", "", "&Jastrow"]

    def block(p1, p2, pq):
        p, q = pq
        return [f"ParticleTypes: {p1} {p2}", "CorrelationFunctionType: Pade",
                "NumberOfParameterTypes: 2",
                f"NumberOfParametersOfEachType: {len(p)} {len(q)}",
                "Parameters: " + " ".join(_fmt(x) for x in list(p) + list(q)),
                "NumberOfConstantTypes: 0", "NumberOfConstantsOfEachType:",
                "Constants:", ""]
    if sys_.n_dn > 0:
        L += block("Electron_up", "Electron_down", sys_.ee_updn)
    if sys_.n_up > 1:
        L += block("Electron_up", "Electron_up", sys_.ee_upup)
    seen = []
    for lab in sys_.labels:
        if lab not in seen:
            seen.append(lab)
            L += block("Electron_up", lab, sys_.en[lab])
    L += ["&Jastrow", ""]
    with open(path, "w") as f:
        f.write("\n".join(L))
