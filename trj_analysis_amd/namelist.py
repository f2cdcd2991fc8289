"""Reader of the reference's Fortran namelists for the structural path.

Restates the parts of src/io/input.f90 the hot path uses:
  /INPUT/      trj_input_file, ndim, nsp, nthread, periodic, ncfs_from_to,
               rdf_sq_cl_dyn_sqw_conf_ord (input.f90:64-67; nthread defaults
               to 64, input.f90:51)
  /INPUT_SP/   sp_types_selected (input.f90:68, 100-102; zeros where unset)
  /INPUT_RDF/  deltar, rcrdf, nrandom (input.f90:69); deltar = -1 sentinel (:143)
  /INPUT_SQ/   qmax, qmin, bsc (input.f90:70); qmax = -1 sentinel (:147),
               bsc defaults to 1.0 per species (:106-107)
The groups are read only when the module flags ask for them (input.f90:144-149):
RDF for flags 1, 3 or 6; S(Q) for flags 2, 3 or 5.  qmin and rcrdf have no
default in the reference; here qmin defaults to qmax (the reference reaches an
unallocated-array kernel launch without it, sq.cuf:110-117) and rcrdf to -1
(half the shortest side, rdf.cuf:55-58).
"""
from __future__ import annotations

import re
from dataclasses import dataclass, field
from typing import Dict, List, Optional


def _parse_scalar(tok: str):
    t = tok.strip()
    if not t:
        return None
    tl = t.lower()
    if tl in (".t.", ".true.", "t", "true"):
        return True
    if tl in (".f.", ".false.", "f", "false"):
        return False
    if (t[0] == t[-1]) and t[0] in "'\"" and len(t) >= 2:
        return t[1:-1]
    try:
        if re.fullmatch(r"[+-]?\d+", t):
            return int(t)
        return float(t.lower().replace("d", "e"))
    except ValueError:
        return t


def _split_values(s: str) -> List:
    toks = re.findall(r"'[^']*'|\"[^\"]*\"|[^\s,]+", s)
    out = []
    for tok in toks:
        m = re.fullmatch(r"(\d+)\*(.*)", tok)
        if m and not tok.startswith(("'", '"')):
            out.extend([_parse_scalar(m.group(2))] * int(m.group(1)))
        else:
            out.append(_parse_scalar(tok))
    return out


def parse_groups(text: str) -> Dict[str, Dict[str, List]]:
    """All namelist groups of a file: {group: {key: [values...]}} (lower case)."""
    lines = [ln.split("!")[0] for ln in text.splitlines()]
    body = "\n".join(lines)
    groups: Dict[str, Dict[str, List]] = {}
    for m in re.finditer(r"&(\w+)(.*?)\n\s*/", body, re.S):
        name = m.group(1).lower()
        content = m.group(2)
        entries: Dict[str, List] = {}
        parts = re.split(r"(\b[A-Za-z_]\w*(?:\(\s*\d+\s*(?::\s*\d+\s*)?\))?)\s*=", content)
        # parts: [pre, key1, val1, key2, val2, ...]
        for k in range(1, len(parts) - 1, 2):
            key = parts[k].strip().lower()
            entries[key] = _split_values(parts[k + 1])
        groups[name] = entries
    return groups


@dataclass
class StructuralInput:
    trj_input_file: str = ""
    ndim: int = 3
    nsp: int = 1
    nthread: int = 64
    periodic: tuple = (True, True, True)
    ncfs_from_to: tuple = (0, 1, 0)
    flags: tuple = (False,) * 7
    deltar: float = -1.0
    rcrdf: float = -1.0
    nrandom: int = 0
    qmax: float = -1.0
    qmin: Optional[float] = None
    bsc: List[float] = field(default_factory=list)
    sp_types_selected: List[int] = field(default_factory=list)
    nslice: int = -1                 # /INPUT/ (common.cuf:55: default -1)
    zslice: List[float] = field(default_factory=list)   # /INPUT_CONF/
    zgrid: Optional[float] = None
    auto_zslice: bool = False        # nslice < 1: one slice set from the first frame

    @property
    def run_conf(self) -> bool:
        return bool(self.flags[5])

    @property
    def slices(self) -> bool:
        """twoDstruc_3D (input.f90:186-189): confinement with RDF or S(Q)."""
        return self.run_conf and (self.run_rdf or self.run_sq)

    @property
    def run_rdf(self) -> bool:
        return bool(self.flags[0])

    @property
    def run_sq(self) -> bool:
        return bool(self.flags[1])


def read_input(text: str) -> StructuralInput:
    """read_input_file (input.f90:78-217) restricted to the structural path."""
    g = parse_groups(text)
    inp = StructuralInput()
    i = g.get("input", {})
    if "trj_input_file" in i:
        inp.trj_input_file = str(i["trj_input_file"][0])
    if "ndim" in i:
        inp.ndim = int(i["ndim"][0])
    if "nsp" in i:
        inp.nsp = int(i["nsp"][0])
    if "nthread" in i:
        inp.nthread = int(i["nthread"][0])
    if "periodic" in i:
        p = [bool(v) for v in i["periodic"]] + [True] * 3
        inp.periodic = tuple(p[:3])
    if "ncfs_from_to" in i:
        v = [int(x) for x in i["ncfs_from_to"]] + [0, 0, 0]
        inp.ncfs_from_to = tuple(v[:3])
    if inp.ncfs_from_to[1] == 0:  # input.f90:137-140
        inp.ncfs_from_to = (inp.ncfs_from_to[0], 1, inp.ncfs_from_to[2])
    if "rdf_sq_cl_dyn_sqw_conf_ord" in i:
        f = [bool(v) for v in i["rdf_sq_cl_dyn_sqw_conf_ord"]] + [False] * 7
        inp.flags = tuple(f[:7])
    if "nslice" in i:
        inp.nslice = int(i["nslice"][0])
    inp.bsc = [1.0] * inp.nsp
    sp = g.get("input_sp", {})
    sel = [int(x) for x in sp.get("sp_types_selected", [])]
    inp.sp_types_selected = (sel + [0] * inp.nsp)[: inp.nsp]
    fl = inp.flags
    if fl[0] or fl[2] or fl[5]:
        r = g.get("input_rdf", {})
        if "deltar" in r:
            inp.deltar = float(r["deltar"][0])
        if "rcrdf" in r:
            inp.rcrdf = float(r["rcrdf"][0])
        if "nrandom" in r:
            inp.nrandom = int(r["nrandom"][0])
    if fl[1] or fl[2] or fl[4]:
        s = g.get("input_sq", {})
        if "qmax" in s:
            inp.qmax = float(s["qmax"][0])
        if "qmin" in s:
            inp.qmin = float(s["qmin"][0])
        if "bsc" in s:
            b = [float(x) for x in s["bsc"]]
            inp.bsc = (b + [1.0] * inp.nsp)[: inp.nsp]
    if inp.qmin is None:
        inp.qmin = inp.qmax
    if fl[5]:  # input.f90:166-191
        if inp.nslice < 1:
            inp.nslice = 1
            inp.auto_zslice = True
        c = g.get("input_conf", {})
        z = [float(x) for x in c.get("zslice", [])]
        inp.zslice = (z + [0.0] * inp.nslice)[: inp.nslice]
        if "zgrid" in c:
            inp.zgrid = float(c["zgrid"][0])
    return inp


def read_input_file(path: str) -> StructuralInput:
    with open(path) as f:
        return read_input(f.read())
