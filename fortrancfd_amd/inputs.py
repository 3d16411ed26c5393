"""Readers for the reference's problem-deck input files, returning the values the library's
set-up calls take (Solver.set_flow_boundary / set_radiation / set_wall_heat / set_ignition).

Each reader follows the Fortran read it replaces: namelist groups (``&name key = value ... /``,
SRC/params.f90 and SRC/mfr.f90:21-38, SRC/ignitionsource.f90:40-70) and list-directed reads that
take the leading value(s) of each line and ignore the trailing comment (SRC/flow_bound.f90:15-57,
SRC/radiation.f90:21-71, SRC/ignitionsource.f90:88-147).  Fortran's ``d`` exponent is accepted.
"""
import re

import numpy as np

_NUM = re.compile(r"^[+-]?(\d+\.?\d*|\.\d+)([dDeE][+-]?\d+)?$")


def _value(tok: str):
    tok = tok.strip().rstrip(",")
    if len(tok) >= 2 and tok[0] == tok[-1] and tok[0] in "'\"":
        return tok[1:-1]
    if tok.lower() in (".true.", "t", ".t."):
        return True
    if tok.lower() in (".false.", "f", ".f."):
        return False
    if _NUM.match(tok):
        v = float(tok.replace("d", "e").replace("D", "e"))
        return int(v) if re.match(r"^[+-]?\d+$", tok) else v
    raise ValueError(f"not a Fortran literal: {tok!r}")


def _strip_comment(line: str) -> str:
    out, quote = [], None
    for ch in line:
        if quote:
            if ch == quote:
                quote = None
        elif ch in "'\"":
            quote = ch
        elif ch == "!":
            break
        out.append(ch)
    return "".join(out)


def read_namelist(path: str, group: str) -> dict:
    """The ``&group`` namelist of a file as {lower-case name: value}."""
    text = open(path).read().splitlines()
    inside, out = False, {}
    for raw in text:
        line = _strip_comment(raw).strip()
        if not inside:
            if line.lower().startswith("&" + group.lower()):
                inside = True
                line = line[len(group) + 1:].strip()
            else:
                continue
        if line.startswith("/") or line.lower().startswith("&end"):
            return out
        for m in re.finditer(r"(\w+)\s*=\s*('[^']*'|\"[^\"]*\"|[^\s,]+)", line):
            out[m.group(1).lower()] = _value(m.group(2))
    if not inside:
        raise KeyError(f"{path}: no &{group} group")
    return out


def _lines(path):
    """Non-empty lines, comments removed."""
    for raw in open(path):
        line = _strip_comment(raw).strip()
        if line:
            yield line


def read_flow_boundary(path: str, names) -> dict:
    """set_flow_boundary (SRC/flow_bound.f90:15-57): inflow T (K), inflow velocity (cm/s), the
    inflow species count and (name, mole fraction) lines, outflow pressure (atm).  Mole fractions of
    species the mechanism lacks are dropped, as the reference's name match drops them."""
    it = _lines(path)
    T_in = float(_value(next(it).split()[0]))
    u_in = float(_value(next(it).split()[0]))
    n = int(_value(next(it).split()[0]))
    x = np.zeros(len(names))
    index = {s: k for k, s in enumerate(names)}
    for _ in range(n):
        sp, val = next(it).split()[:2]
        if sp in index:
            x[index[sp]] = float(_value(val))
    p_out = float(_value(next(it).split()[0]))
    return {"T_in": T_in, "u_in_cm_s": u_in, "x_in": x, "p_out_atm": p_out}


def read_radiation_params(path: str, names):
    """set_radiation_params (SRC/radiation.f90:21-71): one line per radiating species, its name and
    six coefficients.  Returns (0-based species indices, (n, 6) coefficients).  A species the
    mechanism lacks is an error here (the reference leaves its index undefined)."""
    index = {s: k for k, s in enumerate(names)}
    idx, pmac = [], []
    for line in _lines(path):
        tok = line.split()
        if tok[0] not in index:
            raise KeyError(f"{path}: radiating species {tok[0]} is not in the mechanism")
        idx.append(index[tok[0]])
        pmac.append([float(_value(v)) for v in tok[1:7]])
    return np.array(idx, np.int32), np.array(pmac, np.float64).reshape(-1, 6)


def read_wall_temperature(path: str) -> dict:
    """set_wallT_boundary's &mfrwalltemp namelist (SRC/mfr.f90:21-38)."""
    g = read_namelist(path, "mfrwalltemp")
    return {"Nu": float(g["nu"]), "d_mm": float(g["d_mm"]), "Tw_min": float(g["tw_min"]),
            "Tw_max": float(g["tw_max"]), "alpha": float(g["alpha"]), "beta": float(g["beta"])}


def read_igsource(path: str, multiple: bool = False) -> dict:
    """set_igsource's &probEsource namelist (SRC/ignitionsource.f90:40-85; an absent tig_0 is the
    module variable's initial 0) or set_multi_igsource's list (:88-147: the count, then Eig, rig,
    tig, tig_0 per source).  Units as in the files: mJ, um, us, us."""
    if not multiple:
        g = read_namelist(path, "probEsource")
        return {"E_mJ": [float(g["eig"])], "r_um": [float(g["rig"])], "t_us": [float(g["tig"])],
                "t0_us": [float(g.get("tig_0", 0.0))]}
    it = _lines(path)
    n = int(_value(next(it).split()[0]))
    out = {"E_mJ": [], "r_um": [], "t_us": [], "t0_us": []}
    for _ in range(n):
        for key in ("E_mJ", "r_um", "t_us", "t0_us"):
            out[key].append(float(_value(next(it).split()[0])))
    return out
