"""params.in reader: MOLDY's key = value parameter file (src/params_m.F90:187-631).

read_params follows readparam_keyvaluepair: blank lines and lines starting
with '#' are skipped; every other line is "key = value" with a registered key
(case-insensitive; an unknown key or a line without '=' is an error, as the
reference stops with "READ_PARAMS: READ ERROR IN PARAMS.IN"); the value is
read list-directed, so whatever follows the items the key needs (a trailing
"# comment") is ignored.  deltat is converted from fs to model time and
press/pressstep from GPa to model pressure as they are read (:402-404,
:452-455, :578-581); after the file, the box mass becomes nm * boxmass and
bdel2 = deltat^2 / (2 bmass) (:206-208).  Defaults are those of the
simparameters type (:44-160).
"""
from __future__ import annotations

import dataclasses
import re

import numpy as np

ELECTRON = 1.602176487e-19           # src/constants_m.F90:88
ANGSTROM = 1.e-10
TIME_TO_SEC = 1.018050530893617e-014
FS_TO_TIME = 1.0 / (TIME_TO_SEC / 1.0e-15)
GPA_TO_PRESS = 1.0 / ((ELECTRON / ANGSTROM ** 3) / 1.0e9)


class ParamsError(ValueError):
    pass


@dataclasses.dataclass
class Params:
    title1: str = ""
    title2: str = ""
    file_system: str = "system.in"
    file_textout: str = "output.txt"
    file_dumpx1: str = ""
    file_checkpointread: str = ""
    file_checkpointwrite: str = ""
    nlcx: int = 3
    nlcy: int = 3
    nlcz: int = 3
    nm: int = 0
    nspec: int = 0
    dsp: float = 0.0
    rpad: float = 0.0
    deltat: float = 0.0         # model time units
    deltat_fs: float = 0.0      # as written in params.in
    temprq: float = 0.0
    tempsp: float = 0.0
    zlayer: float = 0.0
    pressstep: float = 0.0      # model pressure units
    press: float = 0.0          # model pressure units
    ivol: int = -1
    iquen: int = -1
    iverlet: int = 0
    nsteps: int = 0
    nprint: int = 100
    ntcm: int = 0               # key nbrupdate
    nchkpt: int = -1
    nposav: int = 0
    restart: int = 0
    reset: bool = False
    nose: float = 0.0
    therm_con: float = -1.0
    alternate_quench_md: bool = False
    dumpx1: bool = False
    write_rdf: bool = False
    tjob: float = 0.0
    tfinalise: float = 0.0
    ntc: int = 0
    strx: np.ndarray = dataclasses.field(default_factory=lambda: np.eye(3))  # strx(:,j), column-major
    strainloops: int = 1
    uselookup: bool = False
    nnbrs: int = 0
    pka: int = 0
    pkavx: float = 0.0
    pkavy: float = 0.0
    pkavz: float = 0.0
    epka: float = 0.0
    metropolis: int = -1
    chem_pot: float = 0.0
    metropolis_finetune_on: int = 0
    dialloy_analysis: bool = False
    # derived after reading (:206-208)
    bmass: float = 0.0
    bdel2: float = 0.0


# key -> (attribute, kind, count); kinds: s string, i integer, r real, l logical
_KEYS = {
    "title1": ("title1", "s", 1), "title2": ("title2", "s", 1),
    "file_system": ("file_system", "s", 1), "file_textout": ("file_textout", "s", 1),
    "file_dumpx1": ("file_dumpx1", "s", 1), "boxmass": ("bmass", "r", 1), "deltat": ("deltat", "r", 1),
    "rpad": ("rpad", "r", 1), "nbrupdate": ("ntcm", "i", 1), "strainloops": ("strainloops", "i", 1),
    "pkavx": ("pkavx", "r", 1), "pkavy": ("pkavy", "r", 1), "pkavz": ("pkavz", "r", 1),
    "ntc": ("ntc", "i", 1), "nchkpt": ("nchkpt", "i", 1),
    "alternate_quench_md": ("alternate_quench_md", "l", 1), "temprq": ("temprq", "r", 1),
    "press": ("press", "r", 1), "restart": ("restart", "i", 1), "nsteps": ("nsteps", "i", 1),
    "nprint": ("nprint", "i", 1), "therm_con": ("therm_con", "r", 1), "tjob": ("tjob", "r", 1),
    "tfinalise": ("tfinalise", "r", 1), "iquen": ("iquen", "i", 1), "ivol": ("ivol", "i", 1),
    "iverlet": ("iverlet", "i", 1), "nlcx": ("nlcx", "i", 1), "nlcy": ("nlcy", "i", 1),
    "nlcz": ("nlcz", "i", 1), "dsp": ("dsp", "r", 1), "nnbrs": ("nnbrs", "i", 1),
    "dumpx1": ("dumpx1", "l", 1), "straintensor_row1": ("strx1", "r", 3),
    "straintensor_row2": ("strx2", "r", 3), "straintensor_row3": ("strx3", "r", 3),
    "nm": ("nm", "i", 1), "nspec": ("nspec", "i", 1), "straintensor": ("strx", "r", 9),
    "uselookup": ("uselookup", "l", 1), "file_checkpointread": ("file_checkpointread", "s", 1),
    "file_checkpointwrite": ("file_checkpointwrite", "s", 1), "tempsp": ("tempsp", "r", 1),
    "zlayer": ("zlayer", "r", 1), "write_rdf": ("write_rdf", "l", 1), "pka": ("pka", "i", 1),
    "epka": ("epka", "r", 1), "pressstep": ("pressstep", "r", 1), "nose": ("nose", "r", 1),
    "nposav": ("nposav", "i", 1), "reset": ("reset", "l", 1),
    # key 53 'atomstress' is read into simparam%reset in the reference (:590-591)
    "atomstress": ("reset", "l", 1),
    "metropolis": ("metropolis", "i", 1), "chem_pot": ("chem_pot", "r", 1),
    "metropolis_finetune_on": ("metropolis_finetune_on", "i", 1),
    "dialloy_analysis": ("dialloy_analysis", "l", 1),
}


def _items(text: str):
    """List-directed items: quoted strings, or runs up to a blank or comma; a
    slash ends the record."""
    out, i, n = [], 0, len(text)
    while i < n:
        c = text[i]
        if c in " \t,":
            i += 1
            continue
        if c == "/":
            break
        if c in "\"'":
            j = text.find(c, i + 1)
            if j < 0:
                raise ParamsError("unterminated string: " + text)
            out.append(text[i + 1:j])
            i = j + 1
            continue
        m = re.match(r"[^ \t,/]+", text[i:])
        out.append(m.group(0))
        i += len(m.group(0))
    return out


def _real(tok: str) -> float:
    t = tok.strip().lower().replace("d", "e").replace("q", "e")
    try:
        return float(t)
    except ValueError as e:
        raise ParamsError("not a real: " + tok) from e


def _logical(tok: str) -> bool:
    t = tok.strip().lower().lstrip(".")
    if t[:1] == "t":
        return True
    if t[:1] == "f":
        return False
    raise ParamsError("not a logical: " + tok)


def parse_params(text: str) -> Params:
    p = Params()
    for lineno, raw in enumerate(text.splitlines(), 1):
        line = raw[:200].strip()
        if not line or line.startswith("#"):
            continue
        eq = line.find("=")
        if eq < 0:
            raise ParamsError(f"line {lineno}: malformed pair: {raw}")
        key = line[:eq].strip().lower()
        if key not in _KEYS:
            raise ParamsError(f"line {lineno}: key is not recognised: {key}")
        attr, kind, count = _KEYS[key]
        items = _items(line[eq + 1:])
        if len(items) < count:
            raise ParamsError(f"line {lineno}: {key} needs {count} value(s)")
        items = items[:count]
        if kind == "s":
            val = items[0]
        elif kind == "i":
            try:
                val = int(items[0])
            except ValueError as e:
                raise ParamsError(f"line {lineno}: not an integer: {items[0]}") from e
        elif kind == "l":
            val = _logical(items[0])
        else:
            vals = [_real(t) for t in items]
            val = vals[0] if count == 1 else vals
        if attr == "deltat":
            p.deltat_fs = val
            val = val * FS_TO_TIME
        elif attr in ("press", "pressstep"):
            val = val * GPA_TO_PRESS
        if attr == "strx":
            p.strx = np.array(val, dtype=np.float64).reshape(3, 3, order="F")
        elif attr in ("strx1", "strx2", "strx3"):
            p.strx[:, int(attr[-1]) - 1] = val   # strx(:,j) = row j as written
        else:
            setattr(p, attr, val)
    p.bmass = p.nm * p.bmass
    p.bdel2 = p.deltat * p.deltat / (2.0 * p.bmass) if p.bmass else 0.0
    return p


def read_params(path: str) -> Params:
    with open(path) as fh:
        return parse_params(fh.read())
