"""Reader for the reference's `.ckmf` input format (the parts the walker-
evaluation path needs): &flags, &geometry, &basis, &wavefunction, &Jastrow.

Semantics follow the reference parsers:
  * &flags: keyword / value pairs until '&' (QMCFlags.cpp:24-923);
  * &geometry: `Label Z x y z` per atom (QMCMolecule.cpp:83-113);
  * &basis: per atom `Label nBF maxGauss`, then per basis function
    `nGauss type` and nGauss `alpha coef` lines, coefficients used raw, type
    strings s/px/dxy/... (QMCBasisFunctionCoefficients.cpp:39-127);
  * &wavefunction: restricted (or unrestricted: Alpha then Beta blocks with a
    three-token header each) orbital coefficients, `Alpha Occupation`,
    `Beta Occupation`, `CI Coeffs` (QMCWavefunction.cpp:311-472); the per-spin
    coefficient matrices are the rows of the orbitals any determinant
    occupies, in orbital order, and each determinant lists its columns among
    them (QMCWavefunction.cpp:873-905); several determinants add the packed
    keys n_det, n_act_up/dn, ci_coef, occ_up/dn;
  * &Jastrow: two-body blocks (QMCCorrelationFunctionParameters.cpp:59-279)
    of type Pade (numerator parameters first, PadeCorrelationFunction.cpp:15-51)
    or FixedCuspPade (constants then parameters in numerator and denominator,
    FixedCuspPadeCorrelationFunction.cpp:15-77); link_Jastrow_parameters 1
    copies up-up to down-down and up-nucleus to down-nucleus
    (QMCJastrowParameters.cpp:880-900).

Only all-electron inputs with Cartesian s/p/d functions and Pade-type
two-body Jastrows map onto the kernels; anything else raises
`UnsupportedInput` naming the feature.
"""
from __future__ import annotations

import re
from typing import Dict, List, Tuple

import numpy as np

from .synth import PADE_MAX, _pade_arr, type_powers


class UnsupportedInput(ValueError):
    pass


def _section(tokens: List[str], name: str) -> List[str]:
    """Tokens between `&name` and the next `&` (or `&name` for &Jastrow)."""
    try:
        i = tokens.index("&" + name)
    except ValueError:
        raise UnsupportedInput(f"no &{name} section")
    out = []
    for t in tokens[i + 1:]:
        if t == "&" or (t.startswith("&") and t.lower() != "&" + name.lower()):
            break
        if t == "&" + name:
            break
        out.append(t)
    return out


def _tokens(text: str) -> List[str]:
    return text.split()


def read_flags_text(text: str) -> Dict[str, str]:
    """&flags keyword/value pairs; a keyword token containing '#' starts a
    comment that runs to the end of its line (QMCFlags.cpp:867-877)."""
    toks = []  # (token, line number)
    in_flags = False
    for ln_no, line in enumerate(text.split("\n")):
        for t in line.split():
            if not in_flags:
                if t == "&flags":
                    in_flags = True
                continue
            if t == "&":
                return _pairs(toks)
            toks.append((t, ln_no))
    return _pairs(toks)


def _pairs(toks) -> Dict[str, str]:
    out, i = {}, 0
    while i < len(toks):
        key, ln = toks[i]
        if "#" in key:
            while i < len(toks) and toks[i][1] == ln:
                i += 1
            continue
        if i + 1 < len(toks):
            out[key] = toks[i + 1][0]
        i += 2
    return out


def evaluate_dt(flags: Dict[str, str]) -> float:
    """Time step the reference's evaluate sees right after reading the input:
    dt_equilibration, raised to dt when smaller (QMCFlags.cpp:914-916,
    1282-1303)."""
    dt = float(flags.get("dt", "0.01"))
    dte = float(flags.get("dt_equilibration", "0.01"))
    if dte < 0 or dt > dte:
        dte = dt
    return dte


def _first_upper(s: str) -> str:
    return s[:1].upper() + s[1:].lower()


def read_ckmf(path: str):
    """Returns (packed system dict as synth.System.pack(), flags dict, info)."""
    text = open(path).read()
    tok = _tokens(text)
    flags = read_flags_text(text)
    natoms = int(flags["atoms"])
    norb = int(flags["norbitals"])
    nbf = int(flags["nbasisfunc"])
    ndet = int(flags.get("ndeterminants", "1"))
    tft = flags.get("trial_function_type", "restricted")
    charge = int(flags.get("charge", "0"))
    link = int(flags.get("link_Jastrow_parameters", "1"))
    if int(flags.get("use_three_body_jastrow", "0")):
        raise UnsupportedInput("three-body Jastrow")
    # ---- geometry
    g = _section(tok, "geometry")
    labels, Z, xyz = [], [], []
    for a in range(natoms):
        lab, z, x, y, zz = g[5 * a:5 * a + 5]
        labels.append(_first_upper(lab))
        Z.append(int(float(z)))
        xyz.append([float(x), float(y), float(zz)])
    if "&pseudo" in tok or "&pseudopotential" in tok:
        raise UnsupportedInput("pseudopotentials")
    # ---- basis
    b = _section(tok, "basis")
    pos = 0
    bf_atom, bf_pow, off, alpha, coef = [], [], [0], [], []
    for a in range(natoms):
        pos += 1  # label
        nb_atom = int(b[pos]); pos += 1
        pos += 1  # max gaussians
        for _ in range(nb_atom):
            ng = int(b[pos]); typ = b[pos + 1].lower(); pos += 2
            if len(typ) - 1 > 2 or typ[0] not in "spd":
                raise UnsupportedInput(f"basis function type {typ} (s/p/d only)")
            for _ in range(ng):
                alpha.append(float(b[pos])); coef.append(float(b[pos + 1])); pos += 2
            bf_atom.append(a)
            bf_pow.append(type_powers(typ))
            off.append(len(alpha))
    if len(bf_atom) != nbf:
        raise UnsupportedInput(f"nbasisfunc {nbf} but {len(bf_atom)} functions read")
    # ---- wavefunction
    w = _section(tok, "wavefunction")
    pos = 0
    nrows = norb if tft == "restricted" else 2 * norb

    def take_orbitals(count):
        nonlocal pos
        vals = np.array([float(v) for v in w[pos:pos + count * nbf]])
        pos += count * nbf
        return vals.reshape(count, nbf)
    if tft == "restricted":
        orbs = take_orbitals(norb)
    elif tft == "unrestricted":
        pos += 3
        oa = take_orbitals(norb)
        pos += 3
        ob = take_orbitals(norb)
        orbs = np.vstack([oa, ob])
    else:
        raise UnsupportedInput(f"trial_function_type {tft}")
    # occupations: ndet rows of 0/1 per spin; the beta rows of an unrestricted
    # input index the second orbital block (QMCWavefunction.cpp:364-381)
    pos += 2  # "Alpha Occupation"
    occ_a = np.array([int(v) for v in w[pos:pos + ndet * norb]]).reshape(ndet, norb)
    pos += ndet * norb
    pos += 2  # "Beta Occupation"
    occ_b = np.array([int(v) for v in w[pos:pos + ndet * norb]]).reshape(ndet, norb)
    pos += ndet * norb
    pos += 2  # "CI Coeffs"
    ci = np.zeros(ndet)
    constraint = [-1] * ndet
    factor = [1.0] * ndet
    for d in range(ndet):
        t = w[pos]; pos += 1
        if "c" in t:  # "c <det> <factor>": tied to another coefficient (:386-409)
            c = int(w[pos]); f = float(w[pos + 1]); pos += 2
            constraint[d] = c if constraint[c] == -1 else constraint[c]
            factor[d] = f
        else:
            ci[d] = float(t)
    for d in range(ndet):
        if constraint[d] != -1:
            ci[d] = factor[d] * ci[constraint[d]]
    # per spin: active orbitals = used by any determinant, in orbital order; a
    # determinant's columns are its occupied orbitals in increasing order
    # (QMCWavefunction::makeCoefficients, QMCWavefunction.cpp:873-905)
    def spin_tables(occ, row_offset):
        act = [o for o in range(norb) if occ[:, o].any()]
        index = {o: i for i, o in enumerate(act)}
        cols = [[index[o] for o in range(norb) if occ[d, o]] for d in range(ndet)]
        return act, cols
    act_a, cols_a = spin_tables(occ_a, 0)
    act_b, cols_b = spin_tables(occ_b, nrows - norb)
    n_up, n_dn = len(cols_a[0]), len(cols_b[0])
    if any(len(c) != n_up for c in cols_a) or any(len(c) != n_dn for c in cols_b):
        raise UnsupportedInput("determinants with different electron counts")
    if n_dn > n_up:
        raise UnsupportedInput("more down than up electrons")
    if sum(Z) != n_up + n_dn + charge:
        raise UnsupportedInput("nuclear charge does not match electrons + charge")
    mo_up = np.ascontiguousarray(orbs[act_a])
    mo_dn = (np.ascontiguousarray(orbs[[o + (nrows - norb) for o in act_b]]) if n_dn
             else np.zeros((0, nbf)))
    multi = ndet > 1 or cols_a[0] != list(range(n_up)) or cols_b[0] != list(range(n_dn))
    # ---- Jastrow
    j = _jastrow_tokens(tok)
    funcs = _parse_jastrow(j)
    zero = ([], [])
    ee_updn = funcs.get(("Electron_up", "Electron_down"), zero)
    ee_upup = funcs.get(("Electron_up", "Electron_up"), zero)
    ee_dndn = ee_upup if link else funcs.get(("Electron_down", "Electron_down"), zero)
    en_up, en_dn = [], []
    for lab in labels:
        up = funcs.get(("Electron_up", lab), zero)
        dn = up if link else funcs.get(("Electron_down", lab), zero)
        en_up.append(_pade_arr(up))
        en_dn.append(_pade_arr(dn))
    packed = {
        "n_up": n_up, "n_dn": n_dn, "n_nuc": natoms, "n_basis": nbf, "n_prim": len(alpha),
        "nuc_xyz": np.asarray(xyz, dtype=np.float64),
        "nuc_z": np.asarray(Z, dtype=np.float64),
        "bf_atom": np.asarray(bf_atom, dtype=np.int32),
        "bf_pow": np.asarray(bf_pow, dtype=np.int32).reshape(-1, 3),
        "bf_prim_off": np.asarray(off, dtype=np.int32),
        "prim_alpha": np.asarray(alpha, dtype=np.float64),
        "prim_coef": np.asarray(coef, dtype=np.float64),
        "mo_up": mo_up, "mo_dn": mo_dn,
        "ee_updn": _pade_arr(ee_updn), "ee_upup": _pade_arr(ee_upup),
        "ee_dndn": _pade_arr(ee_dndn),
        "en_up": np.stack(en_up), "en_dn": np.stack(en_dn),
    }
    if multi:
        packed.update({
            "n_det": ndet, "n_act_up": len(act_a), "n_act_dn": len(act_b),
            "ci_coef": np.asarray(ci, dtype=np.float64),
            "occ_up": np.asarray(cols_a, dtype=np.int32).reshape(ndet, n_up),
            "occ_dn": np.asarray(cols_b, dtype=np.int32).reshape(ndet, n_dn)})
    info = {"labels": labels, "ci_coeff": ci,
            "cusp_replacement": int(flags.get("replace_electron_nucleus_cusps", "1"))}
    return packed, flags, info


def _jastrow_tokens(tok: List[str]) -> List[str]:
    idx = [i for i, t in enumerate(tok) if t == "&Jastrow"]
    if not idx:
        return []
    end = idx[1] if len(idx) > 1 else len(tok)
    return tok[idx[0] + 1:end]


def _parse_jastrow(j: List[str]) -> Dict[Tuple[str, str], Tuple[list, list]]:
    out = {}
    pos = 0
    while pos < len(j):
        if j[pos] != "ParticleTypes:":
            pos += 1
            continue
        p1, p2 = _first_upper(j[pos + 1]), _first_upper(j[pos + 2])
        pos += 3
        if p2.startswith("Electron_") and not p1.startswith("Electron_"):
            p1, p2 = p2, p1
        if p1 == "Electron_down" and p2 == "Electron_up":
            p1, p2 = p2, p1
        assert j[pos] == "CorrelationFunctionType:"
        ftype = j[pos + 1]
        pos += 2
        pos += 1  # NumberOfParameterTypes:
        npt = int(j[pos]); pos += 1
        pos += 1  # NumberOfParametersOfEachType:
        per_type = [int(v) for v in j[pos:pos + npt]]; pos += npt
        assert j[pos] == "Parameters:"
        pos += 1
        params = []
        while pos < len(j) and _is_num(j[pos]):
            params.append(float(j[pos])); pos += 1
        nct, consts, per_const = 0, [], []
        if pos < len(j) and j[pos] == "NumberOfConstantTypes:":
            nct = int(j[pos + 1]); pos += 2
            pos += 1  # NumberOfConstantsOfEachType:
            per_const = [int(v) for v in j[pos:pos + nct]]; pos += nct
            assert j[pos] == "Constants:"
            pos += 1
            while pos < len(j) and _is_num(j[pos]):
                consts.append(float(j[pos])); pos += 1
        if ftype == "Pade":
            n0 = per_type[0]
            num, den = params[:n0], params[n0:]
        elif ftype == "FixedCuspPade":
            nnc = len(consts) if len(per_const) <= 1 else per_const[0]
            nnp = len(params) if len(per_type) <= 1 else per_type[0]
            num = consts[:nnc] + params[:nnp]
            den = consts[nnc:] + params[nnp:]
        else:
            raise UnsupportedInput(f"correlation function {ftype} (Pade/FixedCuspPade only)")
        if len(num) > PADE_MAX or len(den) > PADE_MAX:
            raise UnsupportedInput("Pade order above 8")
        out[(p1, p2)] = (num, den)
    return out


_NUM = re.compile(r"^[-+]?(\d+\.?\d*|\.\d+)([eEdD][-+]?\d+)?$")


def _is_num(t: str) -> bool:
    return bool(_NUM.match(t))


def write_with_flags(path_in: str, path_out: str, overrides: Dict[str, str]) -> None:
    """Copy a .ckmf file, replacing (or adding) &flags entries."""
    text = open(path_in).read()
    lines = text.split("\n")
    out, in_flags, seen = [], False, set()
    i = 0
    while i < len(lines):
        ln = lines[i]
        s = ln.strip()
        if s == "&flags":
            in_flags = True
            out.append(ln)
            i += 1
            continue
        if in_flags and s == "&":
            for k, v in overrides.items():
                if k not in seen:
                    out += [k, " " + v]
            in_flags = False
        if in_flags and s in overrides and i + 1 < len(lines):
            out += [s, " " + overrides[s]]
            seen.add(s)
            i += 2
            continue
        out.append(ln)
        i += 1
    open(path_out, "w").write("\n".join(out))
