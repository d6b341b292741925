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
    return parse_namelist_group(text)[1]


def parse_namelist_group(text: str):
    """(group name, {key: value or list}) of one namelist record, keys in
    the record's order."""
    body = text.strip()
    m = re.match(r"&\s*(\w+)(.*)/", body, re.S)
    if not m:
        raise ValueError("not a namelist record: %r" % body[:40])
    group, rest = m.group(1).upper(), m.group(2)
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
    return group, out


@dataclass
class MPSnapshot:
    version: str
    n_head: int
    header: dict = field(default_factory=dict)
    at_ind: np.ndarray = None     # int32 [n_tot, 4]
    at_pos: np.ndarray = None     # float32 [n_tot, 4]
    time_stamp: str = ""
    # further float32 [n_tot, 4] blocks, present as the header says (N_TRAJ >= 1:
    # velocities, N_TRAJ == 2: forces; J_SHELL_OUT == 1: the shells' blocks)
    at_vel: np.ndarray = None
    at_force: np.ndarray = None
    at_pos_s: np.ndarray = None
    at_vel_s: np.ndarray = None
    at_force_s: np.ndarray = None
    groups: list = field(default_factory=list)  # [(name, {key: value})] in record order

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


# the float32 blocks after AT_POS, in file order (mp_lbin56.f90:1490-1547):
# core velocities (N_TRAJ >= 1), core forces (N_TRAJ == 2), then with
# J_SHELL_OUT == 1 the shell positions, velocities and forces likewise
def _blocks(hdr: dict) -> list:
    n_traj = int(hdr.get("N_TRAJ", 0))
    shells = int(hdr.get("J_SHELL_OUT", 0)) == 1
    b = []
    if n_traj >= 1:
        b.append("at_vel")
    if n_traj == 2:
        b.append("at_force")
    if shells:
        b.append("at_pos_s")
        if n_traj >= 1:
            b.append("at_vel_s")
        if n_traj == 2:
            b.append("at_force_s")
    return b


def read_dat(path: str) -> MPSnapshot:
    """Read one MP_TOOLS snapshot: header, species indices, positions and
    the velocity / force / shell blocks the header announces."""
    raw = open(path, "rb").read()
    rb = 4 * L_REC
    if len(raw) % rb:
        raise ValueError(f"{path}: size {len(raw)} is not a multiple of the {rb}-byte record")
    rec = lambda i: raw[(i - 1) * rb:i * rb]  # 1-based like the Fortran REC=
    head = rec(1).decode("latin1").rstrip("\x00 ")
    if not head.startswith("MP_TOOLS"):
        raise ValueError(f"{path}: no MP_TOOLS header record")
    f = head.split()
    version, n_head = f[1], int(f[2])
    hdr, groups = {}, []
    for i in range(2, n_head + 1):
        g, d = parse_namelist_group(rec(i).decode("latin1").rstrip("\x00 "))
        groups.append((g, d))
        hdr.update(d)
    n_tot, n_rec = int(hdr["N_TOT"]), int(hdr["N_REC"])
    if n_rec * L_REC < 4 * n_tot or (n_rec - 1) * L_REC >= 4 * n_tot:
        raise ValueError(f"{path}: N_REC={n_rec} does not match N_TOT={n_tot} at {L_REC} words per record")
    extra = _blocks(hdr)
    if len(raw) < (n_head + (2 + len(extra)) * n_rec) * rb:
        raise ValueError(f"{path}: truncated ({len(raw)} bytes)")
    o = n_head * rb
    ind = np.frombuffer(raw, np.int32, 4 * n_tot, o).reshape(n_tot, 4).copy()
    o += n_rec * rb
    pos = np.frombuffer(raw, np.float32, 4 * n_tot, o).reshape(n_tot, 4).copy()
    s = MPSnapshot(version=version, n_head=n_head, header=hdr, at_ind=ind, at_pos=pos,
                   time_stamp=" ".join(f[3:]), groups=groups)
    for name in extra:
        o += n_rec * rb
        setattr(s, name, np.frombuffer(raw, np.float32, 4 * n_tot, o).reshape(n_tot, 4).copy())
    if int(s.species_counts.sum()) != n_tot:
        raise ValueError(f"{path}: NSUPER_R sums to {s.species_counts.sum()}, N_TOT is {n_tot}")
    return s


# ---------------------------------------------------------------------------
# writer (code/src/mp_lbin56.f90:1436-1480): record 1 is
# dat_source(16) // version(4) // string(40) // '   ' // time_stamp with
# string the list-directed n_head; records 2..n_head each hold one namelist
# group written by gfortran into a 1024-character internal record; every
# record is 4 * L_REC bytes, the part not written left zero.
# ---------------------------------------------------------------------------

def _nml_value(v, lead: bool) -> str:
    """One namelist value as gfortran writes it: integers left-justified in
    11 columns, REAL(4) as G16.9 (left-justified in 16 after a repeat
    count), character with quote delimiters and its full length."""
    from .mptxt import _g16_9
    if isinstance(v, str):
        return '"' + v + '"'
    if isinstance(v, (int, np.integer)) and not isinstance(v, bool):
        return ("%d" % int(v)).ljust(11)
    f = _g16_9(v)
    return f if lead else f.lstrip(" ").ljust(16)


def namelist_record(group: str, items: dict) -> str:
    """gfortran's ``write(internal, nml=group)`` into one record: '&NAME',
    then per object ' ' (record separator) + ' NAME=' + values, each value
    followed by ','; equal consecutive values as ' r*value'; after every
    fifth value of an object two more blanks; finally '  /'."""
    out = ["&" + group.upper()]
    for key, val in items.items():
        out.append("  " + key.upper() + "=")
        vals = val if isinstance(val, list) else [val]
        i = nv = 0
        while i < len(vals):
            j = i
            while j + 1 < len(vals) and type(vals[j + 1]) is type(vals[i]) and _same(vals[j + 1], vals[i]):
                j += 1
            r = j - i + 1
            if r > 1:
                out.append(" %d*" % r + _nml_value(vals[i], lead=False))
            else:
                out.append(_nml_value(vals[i], lead=True))
            out.append(",")
            nv += 1
            if nv % 5 == 0:
                out.append("  ")
            i = j + 1
    out.append("  /")
    return "".join(out)


def _same(a, b) -> bool:
    if isinstance(a, float):
        return np.float32(a).tobytes() == np.float32(b).tobytes()
    return a == b


def _rec(data: bytes) -> bytes:
    rb = 4 * L_REC
    if len(data) > rb:
        raise ValueError("record longer than %d bytes" % rb)
    return data + b"\x00" * (rb - len(data))


def _array_records(a: np.ndarray) -> bytes:
    flat = np.ascontiguousarray(a).reshape(-1).tobytes()
    n_rec = -(-len(flat) // (4 * L_REC))
    return flat + b"\x00" * (n_rec * 4 * L_REC - len(flat))


def dat_bytes(snap: MPSnapshot) -> bytes:
    """The bytes of an MP_TOOLS snapshot file holding snap's header groups,
    AT_IND, AT_POS and the blocks N_TRAJ / J_SHELL_OUT announce, in the
    writer's order (mp_lbin56.f90:1476-1547)."""
    hdr = {}
    for _, d in snap.groups:  # the header as written (the groups it holds)
        hdr.update(d)
    extra = _blocks(hdr)
    for name in extra:
        if getattr(snap, name) is None:
            raise ValueError(f"the header announces {name}, which the snapshot lacks")
    n_head = 1 + len(snap.groups)
    head = ("MP_TOOLS".ljust(16) + snap.version.ljust(4)[:4] + (" " + "%11d" % n_head).ljust(40)
            + "   " + snap.time_stamp.strip()).ljust(L_REC)
    out = [_rec(head.encode("latin1"))]
    for g, d in snap.groups:
        out.append(_rec(namelist_record(g, d).ljust(L_REC).encode("latin1")))
    out.append(_array_records(np.asarray(snap.at_ind, np.int32)))
    out.append(_array_records(np.asarray(snap.at_pos, np.float32)))
    for name in extra:
        out.append(_array_records(np.asarray(getattr(snap, name), np.float32)))
    return b"".join(out)


def write_dat(path: str, snap: MPSnapshot) -> None:
    with open(path, "wb") as f:
        f.write(dat_bytes(snap))


def load_trajectory(paths, device="cuda", pin: bool = True, block: str = "at_pos"):
    """Read MP_TOOLS snapshots (one per frame) into one device tensor
    float32 [T, n_tot, 4] (x, y, z, charge; species-major atom order), the
    layout the S(Q,w) engine takes; block "at_vel" (or another of the
    snapshot's blocks) gives the velocities mpsq_dos takes.  Frames are
    staged through one pinned host buffer and copied with one transfer.  All
    files must hold the same atoms (N_TOT and NSUPER_R)."""
    import torch

    paths = list(paths)
    if not paths:
        raise ValueError("no snapshot files")
    first = read_dat(paths[0])
    n_tot, counts = first.n_tot, first.species_counts
    host = torch.empty((len(paths), n_tot, 4), dtype=torch.float32,
                       pin_memory=bool(pin and torch.cuda.is_available()))
    hv = host.numpy()

    def take(s, p):
        a = getattr(s, block)
        if a is None:
            raise ValueError(f"{p}: no {block} block (N_TRAJ {s.header.get('N_TRAJ')})")
        return a

    hv[0] = take(first, paths[0])
    for k, p in enumerate(paths[1:], start=1):
        s = read_dat(p)
        if s.n_tot != n_tot or not np.array_equal(s.species_counts, counts):
            raise ValueError(f"{p}: atoms differ from {paths[0]}")
        hv[k] = take(s, p)
    return host.to(device, non_blocking=False), first

