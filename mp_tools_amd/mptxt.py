"""Text output of MP_TOOLS maps: the mp_sqom ``.txt`` table.

The reference writes its S(Q) / S(Q,w) map with Fortran I/O
(code/src/mp_sqom56.f90:1994-2045): a header of list-directed
(``write(3,*)``) and explicitly formatted records, one blank record, then
one record per row i of ``cs_plot(i, j)`` in the format ``(200(1x,5e10.3))``.
This module restates both the record layout and the two Fortran edit
descriptors involved, as gfortran renders them:

* list-directed output: a record starts with one blank; a separator blank
  precedes every item except a character item that follows a character
  item; REAL(4) items are written as ``G16.9E2`` (9 significant digits,
  fixed form for 0.1 <= |x| < 1e9 with four trailing blanks, else
  ``1PE16.8E2``), default INTEGER items as ``I11``;
* ``e10.3``: ``0.ddd`` mantissa, two-digit signed exponent, right-justified
  in 10 columns.

Values are the REAL(4) numbers the reference stores (``cs_plot`` is
real(4), mp_sqom56.f90:132): FP64 maps are rounded to FP32 before they are
formatted.
"""
from __future__ import annotations

import io

import numpy as np

__all__ = ["ld_record", "e10_3", "map_rows", "cs_plot_total", "sqom_table", "write_sqom_table"]


class _R4(float):
    """Marks a list-directed item as REAL(4)."""


def _g16_9(x: float) -> str:
    """gfortran list-directed REAL(4) field (without the separator)."""
    x = float(np.float32(x))
    if x == 0.0:
        return "0.00000000".rjust(12) + " " * 4
    s = "%.8e" % x  # 9 significant digits, correctly rounded
    e = int(s[s.index("e") + 1:])
    if -1 <= e <= 8:  # 0.1 <= |x| < 1e9 after rounding: F(12).(8-e) + 4 blanks
        return ("%.*f" % (8 - e, x)).rjust(12) + " " * 4
    m, ex = s.split("e")
    return (m + "E" + ("%+03d" % int(ex))).upper().rjust(16)


def ld_record(*items) -> str:
    """One list-directed output record (``write(u,*) items``): str items are
    character, int items default integer, other numbers REAL(4); sequences
    are expanded element by element."""
    flat = []
    for it in items:
        if isinstance(it, (list, tuple, np.ndarray)):
            flat.extend(np.asarray(it).tolist() if isinstance(it, np.ndarray) else list(it))
        else:
            flat.append(it)
    if not flat:
        return ""
    out = []
    prev_char = None
    for it in flat:
        is_char = isinstance(it, str)
        if prev_char is None:
            out.append(" ")
        elif not (is_char and prev_char):
            out.append(" ")
        if is_char:
            out.append(it)
        elif isinstance(it, (bool, np.bool_)):
            raise TypeError("logical items are not written by this table")
        elif isinstance(it, (int, np.integer)):
            out.append("%11d" % int(it))
        else:
            out.append(_g16_9(it))
        prev_char = is_char
    return "".join(out)


def e10_3(x: float) -> str:
    """Fortran ``e10.3`` of one REAL(4) value: ``0.ddd`` mantissa."""
    x = float(np.float32(x))
    if x == 0.0:
        return " 0.000E+00"
    if not np.isfinite(x):
        return ("NaN" if np.isnan(x) else ("-Infinity" if x < 0 else "Infinity")).rjust(10)
    s = "%.2e" % abs(x)  # d.dd e+xx, 3 significant digits correctly rounded
    mant, ex = s.split("e")
    digits = mant.replace(".", "")
    ex = int(ex) + 1  # 0.ddd * 10^(ex)
    es = ("E%+03d" % ex) if abs(ex) <= 99 else ("%+04d" % ex)
    body = ("-" if x < 0 else "") + "0." + digits + es
    return body.rjust(10)


def map_rows(cs_plot: np.ndarray) -> list:
    """Records of ``write(3,'(200(1x,5e10.3))') (cs_plot(i,j),j=1,n)`` for
    every row i: groups of five values, each group preceded by one blank
    (at most 200 groups per record, i.e. 1000 values)."""
    a = np.asarray(cs_plot, dtype=np.float32)
    if a.ndim != 2:
        raise ValueError("cs_plot must be 2-D (rows i, columns j)")
    if a.shape[1] > 1000:
        raise ValueError("(200(1x,5e10.3)) holds at most 1000 values per record")
    rows = []
    for i in range(a.shape[0]):
        parts = []
        for g in range(0, a.shape[1], 5):
            parts.append(" " + "".join(e10_3(v) for v in a[i, g:g + 5]))
        rows.append("".join(parts))
    return rows


def cs_plot_total(s_flat: np.ndarray, n_qx: int, n_qy: int) -> np.ndarray:
    """``cs_plot(i, j)`` (i along e1 / k1, j along e2 / k2) from a map in
    CMCL order with k1 fastest (the layout of the amplitudes and of S(Q)):
    element (i, j) is s_flat[j * n_qx + i] (mp_sqom56.f90:1154-1155, 1384)."""
    s = np.asarray(s_flat).reshape(n_qy, n_qx)
    return np.ascontiguousarray(s.T).astype(np.float32)


def _f32(x):
    return np.float32(x)


def _grid_last(lo, hi, n):
    """qx_min + (n_qx - 1.) * (qx_max - qx_min) / (n_qx) in REAL(4)."""
    lo, hi = _f32(lo), _f32(hi)
    return lo + (_f32(n) - _f32(1.0)) * (hi - lo) / _f32(n)


def sqom_table(cs_plot, *, mp_tool="MP_SQOM 1.56", subst_name="", shell="", file_dat="",
               t0=0.0, t_step=0.0, n_row=(1, 1, 1), temp=0.0, at_names=(), at_occup=(),
               weight_scheme="Unit weights", b_coh=(), at_mask=None, atc=(0, 0), t_width=0.0,
               e1=(1.0, 0.0, 0.0), e2=(0.0, 1.0, 0.0), q_center=(0.0, 0.0, 0.0), total=True,
               f_plot=0.0, f_width=0.0, qx_range=(0.0, 1.0), qy_range=(0.0, 1.0), disp=None) -> str:
    """The text of mp_sqom's map file (mp_sqom56.f90:1994-2045).

    ``total`` selects the header of the total-scattering map (abs(j_plot)==1)
    or of an E=const map at energy ``f_plot``; ``qx_range`` / ``qy_range``
    are (min, max) of the plot axes (the record gives min and the last pixel
    min + (n-1)(max-min)/n).  With ``disp`` (a dict with q1_3d, q2_3d, n_q1x,
    n_q2x, n_q1y, n_q2y, q1_x, q2_x, q1_y, q2_y, f_min, f_max) the header of
    the E(Q) dispersion cut (j_disp==1) is written instead; then rows are Q
    points and columns frequencies.
    """
    a = np.asarray(cs_plot, dtype=np.float32)
    n_atom = len(at_names)
    at_mask = [1] * n_atom if at_mask is None else list(at_mask)
    rec = []
    rec.append(ld_record("*****    ", mp_tool.strip(), ": total scattering cross-section     *****"))
    rec.append(ld_record())
    rec.append(ld_record("Substance:    ", subst_name.rstrip(), " ", shell.rstrip()))
    rec.append(ld_record("Input files:  ", "./data/", file_dat.rstrip()))
    rec.append(ld_record("  t0,t_step,n_row(3),n_atom:", _R4(t0), _R4(t_step), [int(v) for v in n_row], int(n_atom)))
    rec.append(ld_record("  temperature [K]", _R4(temp)))
    # '(" Atoms:        ",20(a,3x))' with character(4) names; the 3x after
    # the last name only positions, it writes nothing
    rec.append(" Atoms:        " + "   ".join(str(n)[:4].ljust(4) for n in at_names))
    rec.append(" Occupations:" + "".join("%7.3f" % _f32(v) for v in at_occup))
    rec.append(" " + weight_scheme.rstrip() + " b_coh:      " + "".join("%7.3f" % _f32(v) for v in b_coh))
    rec.append(" Atom masks: " + ("%3d" % at_mask[0] if at_mask else "") + "".join("%7d" % m for m in at_mask[1:]))
    rec.append(" Atom pair correlation: " + "".join("%3d" % int(v) for v in atc))
    rec.append(ld_record("Time-integration (BN window FWHM) [ps]:", _R4(t_width)))
    rec.append(" Q-plane: e1 = [" + "".join("%6.2f" % _f32(v) for v in e1) + "], e2 = ["
               + "".join("%6.2f" % _f32(v) for v in e2) + "]")
    rec.append(ld_record("Q center:", [_R4(v) for v in q_center]))
    if disp is None:
        if total:
            rec.append(ld_record("Total scattering (infinite energy width)", _R4(0.0), _R4(9999.0)))
        else:
            rec.append(ld_record("Energy transfer & resolution [THz]:", _R4(f_plot), _R4(f_width)))
        nx, ny = a.shape
        rec.append(ld_record("Plot size X (lines):  ", nx, _R4(qx_range[0]), _R4(_grid_last(qx_range[0], qx_range[1], nx))))
        rec.append(ld_record("Plot size Y (columns):", ny, _R4(qy_range[0]), _R4(_grid_last(qy_range[0], qy_range[1], ny))))
    else:
        d = disp
        nq, nf = a.shape
        rec.append(ld_record("Energy resolution [THz]", _R4(f_width)))
        rec.append(ld_record("Dispersion range [rlu]:", [_R4(v) for v in d["q1_3d"]], "  ", [_R4(v) for v in d["q2_3d"]]))
        if (d["n_q2x"] - d["n_q1x"]) >= (d["n_q2y"] - d["n_q1y"]):
            lo, hi = d["q1_x"], d["q2_x"]
        else:
            lo, hi = d["q1_y"], d["q2_y"]
        rec.append(ld_record("Plot size X (lines, Q):  ", nq, _R4(lo), _R4(_grid_last(lo, hi, nq))))
        rec.append(ld_record("Plot size Y (columns, f):", nf, _R4(d["f_min"]), _R4(d["f_max"])))
    rec.append(ld_record())
    rec.extend(map_rows(a))
    buf = io.StringIO()
    for r in rec:
        buf.write(r + "\n")
    return buf.getvalue()


def write_sqom_table(path: str, cs_plot, **header) -> None:
    """Write mp_sqom's map file (``sqom_table``) to ``path``; ``cs_plot`` is
    the map as rows i (e1) by columns j (e2), e.g. ``cs_plot_total(S_tot,
    n_qx, n_qy)`` for the total map or the same of one energy's S(Q, w)."""
    with open(path, "w") as f:
        f.write(sqom_table(cs_plot, **header))


def _fixed(x: float, w: int, d: int) -> str:
    """Fortran Fw.d of a REAL(4) value (asterisks when it does not fit)."""
    t = "%*.*f" % (w, d, float(np.float32(x)))
    return t if len(t) <= w else "*" * w


def dos_frequencies(n_freq: int, t_step: float) -> np.ndarray:
    """The frequency column of mp_dos's table: f_max = 0.5 / t_step and
    ff(j) = f_min + (j - 1) (f_max - f_min) / real(n_freq - 1), f_min = 0, in
    REAL(4) (mp_dos56.f90:474, 606-607, 620-622) -- the plot axis, which
    runs to the Nyquist frequency over the n_freq rows."""
    f32 = np.float32
    f_max = f32(f32(0.5) / f32(t_step))
    if n_freq <= 1:
        return np.zeros(max(n_freq, 0), np.float32)
    j = np.arange(n_freq, dtype=np.float32)
    return ((j * f_max).astype(np.float32) / f32(n_freq - 1)).astype(np.float32)


def dos_table(cs_out, *, t0: float, t_step: float, n_row, temp: float, at_names, occ, weights,
              file_dat_t0: str = "", file_dat: str = "", mp_tool: str = "MP_DOS 1.56", shells: bool = False) -> str:
    """The text of mp_dos's output file (mp_dos56.f90:723-736): header records
    (tool, input files, t0 / t_step / n_row / n_atom, temperature, atoms,
    occupations, weights, column names) and one row per frequency,
    (1x, f6.3, 2x, f9.5, 2x, (1x, f9.5)), of the frequency, the total and the
    partial DOS.  ``cs_out`` is [n_freq, nsp + 1] as ``sqw.dos_output``
    returns it (partials, then the total)."""
    cs = np.asarray(cs_out, np.float64)
    nsp = len(at_names)
    if cs.ndim != 2 or cs.shape[1] != nsp + 1:
        raise ValueError("cs_out must be [n_freq, nsp + 1]")
    names = [str(n).ljust(4)[:4] for n in at_names]
    rec = [ld_record("*****    ", mp_tool.strip(), ": total and partial densities of vibrational states     *****"),
           ld_record()]
    if shells:
        rec.append(ld_record("ATTENTION: the DOS corresponds to atom SHELLS!"))
    rec += [ld_record("Input files:  ", file_dat_t0.strip(), " to ", file_dat.strip()),
            ld_record("  t0,t_step,n_row,n_atom:", _f32(t0), _f32(t_step), [int(v) for v in n_row], int(nsp)),
            ld_record("  temperature [K]", _f32(temp)),
            ld_record(),
            " Atoms:      " + "   ".join(names),  # (a, 3x): the last 3x writes nothing
            " Occupations:" + "".join(_fixed(v, 7, 3) for v in occ),
            " Weights:    " + "".join(_fixed(v, 7, 3) for v in weights),
            ld_record(),
            " f [THz] " + "   Total  " + "".join(n.rjust(9) for n in names)]
    ff = dos_frequencies(cs.shape[0], t_step)
    for i in range(cs.shape[0]):
        rec.append(" " + _fixed(ff[i], 6, 3) + "  " + _fixed(cs[i, nsp], 9, 5) + "  "
                   + "".join(" " + _fixed(cs[i, j], 9, 5) for j in range(nsp)))
    return "\n".join(rec) + "\n"


def dos_file_name(file_master: str, nfile_min: int, nfile: int, jfile: int = 1) -> str:
    """mp_dos's output name (mp_dos56.f90:683-698): master_dos_<first>_<last>_<jj>.txt,
    the frame numbers as i4.4 below 10000 (i8, left-adjusted, above)."""
    def num(v):
        return "%04d" % v if v <= 9999 else str(int(v))
    return f"{file_master.strip()}_dos_{num(nfile_min)}_{num(nfile)}_{jfile:02d}.txt"


def write_dos_table(path: str, cs_out, **header) -> None:
    """Write ``dos_table(cs_out, **header)`` to ``path``."""
    with open(path, "w", encoding="utf-8") as f:
        f.write(dos_table(cs_out, **header))

