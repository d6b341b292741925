"""Reader of MP_TOOLS binary snapshot files (.dat).

Format (code/docs/1_mp_data_doc_154.txt:65-113; read as in
code/src/mp_sql56.f90:230-283 (header) and :913-945 (arrays)): a
direct-access file of fixed records of l_rec 32-bit words (l_rec = 1024,
4096-byte records, in MP_TOOLS 1.56).

  record 1            text "MP_TOOLS <version> <n_head> <time stamp>"
  records 2..n_head   text Fortran namelists DATA_HEADER_1, _2 (and _3)
  n_rec records       AT_IND(4, N_TOT) int32: cell indices (1:3), species (4)
  n_rec records       AT_POS(4, N_TOT) float32: position (1:3), charge (4)
  (then velocities / forces / shells when present; not read here)

The last record of each array is zero-padded.  Atoms are stored species by
species (NSUPER_R atoms of species 1, then species 2, ...).
"""
from __future__ import annotations

import re
from dataclasses import dataclass, field

import numpy as np

L_REC = 1024  # 32-bit words per record


def _fortran_value(tok: str):
    tok = tok.strip()
    if tok.startswith('"') or tok.startswith("'"):
        return tok[1:-1]
    try:
        if re.fullmatch(r"[+-]?\d+", tok):
            return int(tok)
        return float(tok.replace("D", "E").replace("d", "e"))
    except ValueError:
        return tok


def parse_namelist(text: str) -> dict:
    """Parse one Fortran namelist record ("&NAME  K=v, K2= 3*1.0, ... /") into
    {key: value or list}; repeat counts (3*x) are expanded."""
    body = text.strip()
    m = re.match(r"&\s*(\w+)(.*)/", body, re.S)
    if not m:
        raise ValueError("not a namelist record: %r" % body[:40])
    rest = m.group(2)
    # split at KEY= boundaries, keeping quoted strings intact
    parts = re.split(r"(\b[A-Z_][A-Z0-9_]*\s*=)", rest)
    out = {}
    for i in range(1, len(parts), 2):
        key = parts[i].rstrip("= ").strip().upper()
        raw = parts[i + 1]
        toks = re.findall(r'"[^"]*"|\'[^\']*\'|[^,\s]+', raw)
        vals = []
        for t in toks:
            rm = re.fullmatch(r"(\d+)\*(.*)", t)
            if rm and not t.startswith('"'):
                vals.extend([_fortran_value(rm.group(2))] * int(rm.group(1)))
            else:
                vals.append(_fortran_value(t))
        out[key] = vals[0] if len(vals) == 1 else vals
    return out


@dataclass
class MPSnapshot:
    version: str
    n_head: int
    header: dict = field(default_factory=dict)
    at_ind: np.ndarray = None     # int32 [n_tot, 4]
    at_pos: np.ndarray = None     # float32 [n_tot, 4]

    @property
    def n_tot(self) -> int:
        return int(self.header["N_TOT"])

    @property
    def species_counts(self) -> np.ndarray:
        return np.atleast_1d(np.asarray(self.header["NSUPER_R"], np.int64))

    @property
    def species_names(self) -> list:
        v = self.header["AT_NAME_OUT"]
        return [s.strip() for s in (v if isinstance(v, list) else [v])]

    @property
    def occupancies(self) -> np.ndarray:
        o = np.atleast_1d(np.asarray(self.header["AT_OCCUP_R"], np.float64))
        return o / o.sum()  # normalised as at_occup_r/sum(at_occup_r) (mp_sql56.f90:292)

    @property
    def a_par(self) -> np.ndarray:
        return np.asarray(self.header["A_PAR"], np.float64)

    @property
    def a_cell_inv(self) -> np.ndarray:
        """3x3 (column = reciprocal translation e_r); identity / a_par for
        headers without DATA_HEADER_3 (mp_sql56.f90:645-652)."""
        if "A_CELL_INV" in self.header:
            return np.asarray(self.header["A_CELL_INV"], np.float64).reshape(3, 3, order="F")
        return np.diag(1.0 / self.a_par)


def read_dat(path: str) -> MPSnapshot:
    """Read one MP_TOOLS snapshot (header, species indices, positions)."""
    raw = open(path, "rb").read()
    rb = 4 * L_REC
    if len(raw) % rb:
        raise ValueError(f"{path}: size {len(raw)} is not a multiple of the {rb}-byte record")
    rec = lambda i: raw[(i - 1) * rb:i * rb]  # 1-based like the Fortran REC=
    head = rec(1).decode("latin1")
    if not head.startswith("MP_TOOLS"):
        raise ValueError(f"{path}: no MP_TOOLS header record")
    f = head.split()
    version, n_head = f[1], int(f[2])
    hdr = {}
    for i in range(2, n_head + 1):
        hdr.update(parse_namelist(rec(i).decode("latin1").rstrip("\x00 ")))
    n_tot, n_rec = int(hdr["N_TOT"]), int(hdr["N_REC"])
    if n_rec * L_REC < 4 * n_tot or (n_rec - 1) * L_REC >= 4 * n_tot:
        raise ValueError(f"{path}: N_REC={n_rec} does not match N_TOT={n_tot} at {L_REC} words per record")
    if len(raw) < (n_head + 2 * n_rec) * rb:
        raise ValueError(f"{path}: truncated ({len(raw)} bytes)")
    o = n_head * rb
    ind = np.frombuffer(raw, np.int32, 4 * n_tot, o).reshape(n_tot, 4).copy()
    o += n_rec * rb
    pos = np.frombuffer(raw, np.float32, 4 * n_tot, o).reshape(n_tot, 4).copy()
    s = MPSnapshot(version=version, n_head=n_head, header=hdr, at_ind=ind, at_pos=pos)
    if int(s.species_counts.sum()) != n_tot:
        raise ValueError(f"{path}: NSUPER_R sums to {s.species_counts.sum()}, N_TOT is {n_tot}")
    return s
