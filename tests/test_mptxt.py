"""mp_sqom map-file writer (mp_sqom56.f90:1994-2045) against the reference's
own BZO tables (tests/golden/bzo, extracted by make_bzo_txt.py) and closed
forms of the Fortran edit descriptors.

The BZO maps' trajectory is not in the reference tree, so the tables pin the
text layout only: the header is rebuilt from the values it prints and must
match record by record, and the map records are parsed back to REAL(4) and
must be rewritten unchanged (rows, groups of five, e10.3)."""
import json
import os

import numpy as np
import pytest

from mp_tools_amd import mptxt

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.fixture(scope="module")
def tables():
    return json.load(open(os.path.join(HERE, "golden", "bzo", "bzo_sq_tables.json")))


# the values the BZO tables print (BZO200d.par / the example's mp_sqom run)
BZO = dict(mp_tool="MP_SQOM 1.56", subst_name="BZO200d_demo", shell="", file_dat="bzo200d_n0001....0020.dat",
           t0=np.float32("64.0400009"), t_step=np.float32("4.00009155E-02"), n_row=(32, 32, 32),
           temp=np.float32("199.586273"), at_names=("Ba", "Zr", "O1", "O2", "O3"), at_occup=(1, 1, 1, 1, 1),
           weight_scheme="Unit weights", b_coh=(0.507, 0.716, 0.580, 0.580, 0.580), at_mask=(1, 1, 1, 1, 1),
           atc=(0, 0), t_width=np.float32("0.200004578"), e1=(1, 0, 0), e2=(0, 1, 0), q_center=(2, 2, 0),
           qx_range=(-0.5, 4.5), qy_range=(-0.5, 4.5))


def _parse_rows(rows):
    return np.array([[np.float32(v) for v in r.split()] for r in rows], dtype=np.float32)


@pytest.mark.parametrize("tag", ["07", "11"])
def test_bzo_header_records(tables, tag):
    t = tables[tag]
    kw = dict(BZO)
    if tag == "11":  # E = 0 map at the example's resolution
        kw.update(total=False, f_plot=0.0, f_width=np.float32("4.99988556"))
    cs = np.zeros((t["n_rows"], 160), np.float32)
    text = mptxt.sqom_table(cs, **kw).split("\n")
    for k, (got, want) in enumerate(zip(text[:18], t["header"])):
        # the example's substance record carries the untrimmed 40-character
        # name (trailing blanks); every other record matches to the byte
        if k == 2:
            assert got.rstrip() == want.rstrip(), k
        else:
            assert got == want, (k, got, want)


@pytest.mark.parametrize("tag", ["07", "11"])
def test_bzo_map_records_round_trip(tables, tag):
    t = tables[tag]
    for rows in (t["first_rows"], t["last_rows"]):
        a = _parse_rows(rows)
        assert a.shape == (3, 160)
        assert mptxt.map_rows(a) == rows
    assert t["row_lengths"] == [len(mptxt.map_rows(np.ones((1, 160), np.float32))[0])]


def test_e10_3_closed_forms():
    assert mptxt.e10_3(0.0) == " 0.000E+00"
    assert mptxt.e10_3(1.0) == " 0.100E+01"
    assert mptxt.e10_3(-1.0) == "-0.100E+01"
    assert mptxt.e10_3(0.2425e-3) in (" 0.242E-03", " 0.243E-03")  # tie in decimal, not in binary
    assert mptxt.e10_3(0.24251e-3) == " 0.243E-03"
    assert mptxt.e10_3(9.9996) == " 0.100E+02"  # rounding carries into the exponent
    assert mptxt.e10_3(123456.0) == " 0.123E+06"
    assert mptxt.e10_3(1.4e-45) == " 0.140E-44"  # smallest REAL(4) denormal
    assert mptxt.e10_3(3.0e38) == " 0.300E+39"


def test_list_directed_closed_forms():
    assert mptxt.ld_record() == ""
    assert mptxt.ld_record("a", "b") == " ab"  # no separator between character items
    assert mptxt.ld_record("x", 7) == " x" + " " + " " * 10 + "7"  # separator, then I11
    assert mptxt.ld_record(np.float32(0.5)) == "  0.500000000    "
    assert mptxt.ld_record(np.float32(-0.5)) == " -0.500000000    "
    assert mptxt.ld_record(np.float32(1.0e9)) == "   1.00000000E+09"  # E form at 1e9
    # E form below 0.1; REAL(4) 0.09 is 0.0900000035762787...
    assert mptxt.ld_record(np.float32(0.09)) == "   9.00000036E-02"
    assert mptxt.ld_record(np.float32(999999999.0)) == "   1.00000000E+09"  # rounds to 1e9: E form


def test_cs_plot_orientation():
    # S(Q) in CMCL order, k1 fastest: rows i are k1, columns j are k2
    n_qx, n_qy = 4, 3
    s = np.arange(n_qx * n_qy, dtype=np.float64)
    cs = mptxt.cs_plot_total(s, n_qx, n_qy)
    assert cs.shape == (4, 3)
    for i in range(n_qx):
        for j in range(n_qy):
            assert cs[i, j] == s[j * n_qx + i]


def test_full_table_shape():
    cs = np.random.default_rng(1).random((7, 12)).astype(np.float32)
    text = mptxt.sqom_table(cs, **BZO).split("\n")
    assert text[17] == "" and text[-1] == ""
    rows = text[18:-1]
    assert len(rows) == 7
    assert all(len(r) == 2 * 51 + 1 + 2 * 10 for r in rows)  # two full groups, one of two values
    back = _parse_rows(rows)
    assert np.allclose(back, cs, rtol=5e-3)


def test_qplane_reproduces_the_bzo_map_axes(tables):
    """The Q-plane set-up (mp_sqom56.f90:800-857, 884-907) for the BZO run
    (32^3 cells, 5 Brillouin zones, (hk0) plane, Q centre (2, 2, 0)) gives the
    pixel counts and axes the reference's map files print."""
    from mp_tools_amd import sqw
    g = sqw.qplane((32, 32, 32), 5, (1, 0, 0), (0, 1, 0), (2, 2, 0))
    assert (g["n_qx"], g["n_qy"]) == (160, 160)
    assert mptxt.ld_record("Plot size X (lines):  ", g["n_qx"], mptxt._R4(g["qx_range"][0]),
                           mptxt._R4(mptxt._grid_last(*g["qx_range"], g["n_qx"]))) == tables["07"]["header"][15]
    assert mptxt.ld_record("Plot size Y (columns):", g["n_qy"], mptxt._R4(g["qy_range"][0]),
                           mptxt._R4(mptxt._grid_last(*g["qy_range"], g["n_qy"]))) == tables["07"]["header"][16]
    np.testing.assert_allclose(g["tpe"][:2], (2 * np.pi / 32, 2 * np.pi / 32), rtol=1e-7)


def test_qplane_closed_forms():
    from mp_tools_amd import sqw
    # (hhl): e1 = [110]/sqrt2, e2 = [001]; e1 . (n_row box) = 32 sqrt2 cells
    g = sqw.qplane((32, 32, 32), 4, (1, 1, 0), (0, 0, 1), (1, 1, 0))
    np.testing.assert_allclose(g["e1"], np.array([1, 1, 0]) / np.sqrt(2), rtol=1e-7)
    np.testing.assert_allclose(g["ev"], np.array([1, -1, 0]) / np.sqrt(2), rtol=1e-7)
    n = int(4 * 32 * np.sqrt(2))
    assert g["n_qx"] == n + (n % 2) and g["n_qy"] == 4 * 32
    # qx centre: (e1 . q) / |e1_in| = (1 + 1) / sqrt2 / sqrt2 = 1
    assert abs(float(g["qx_range"][0]) - (1.0 - 2.0)) < 1e-6 and abs(float(g["qy_range"][0]) + 2.0) < 1e-6
    # one bulk box of 40 A (nsuper 1): pixels = bz_n / tpe, bz_n in 1/A
    g = sqw.qplane((1, 1, 1), 2.0, q_center=(0, 0, 0), a_par=(40.0, 40.0, 40.0), nsuper=1)
    assert g["n_qx"] == int(np.float32(2.0) / np.float32(2 * np.pi / 40.0)) + (int(np.float32(2.0) / np.float32(2 * np.pi / 40.0)) % 2)
    assert g["engine_kw"]["axes"].shape == (3, 3)


def test_dos_table_against_the_reference_bzo_file():
    """mp_dos's table (mp_dos56.f90:723-736) for the reference's BZO example
    (bzo200d_dos_0001_0020_01.txt, 20 snapshots, t_step 0.04 ps): the header
    records come out byte for byte, and the row count and frequency field of
    every row match -- the frequency count int(0.5 (nfile - 1)) + 1 and the
    REAL(4) axis to 0.5 / t_step (5.555, not 5.556, at t_step = 0.0400009155).
    The file was written by MP_DOS55, whose value fields had three decimals,
    so the rows' value fields are not compared."""
    import json
    from mp_tools_amd import mptxt, sqw
    gold = json.load(open(os.path.join(HERE, "golden", "bzo", "bzo_dos_table.json"), encoding="utf-8"))
    t_step = np.float32(4.00009155e-02)
    n_freq = sqw.dos_n_freq(20, float(t_step))
    assert n_freq == gold["n_rows"]
    cs = np.zeros((n_freq, 6))
    txt = mptxt.dos_table(cs, t0=64.04, t_step=t_step, n_row=(32, 32, 32), temp=np.float32(199.586273),
                          at_names=["Ba", "Zr", "O1", "O2", "O3"], occ=[1.0] * 5,
                          weights=[25.705, 51.266, 33.675, 33.675, 33.675],
                          file_dat_t0="./data/bzo200d_n0001.dat", file_dat="./data/bzo200d_n0020.dat",
                          mp_tool="MP_DOS55")
    lines = txt.split("\n")[:-1]
    assert lines[:len(gold["header"])] == gold["header"]
    rows = lines[len(gold["header"]):]
    assert len(rows) == gold["n_rows"]
    assert [r[:7] for r in rows] == gold["freq_fields"]
    assert mptxt.dos_file_name("bzo200d", 1, 20, 1) == "bzo200d_dos_0001_0020_01.txt"
    # mp_dos56's own row format: (1x, f6.3, 2x, f9.5, 2x, (1x, f9.5))
    cs[3, :] = [0.1, 0.2, 0.3, 0.4, 0.5, 1.5]
    r3 = mptxt.dos_table(cs, t0=0, t_step=t_step, n_row=(1, 1, 1), temp=0, at_names=["Ba", "Zr", "O1", "O2", "O3"],
                         occ=[1] * 5, weights=[1] * 5).split("\n")[11 + 3]
    assert r3 == " " + " 4.167" + "  " + "  1.50000" + "  " + "".join(" %9.5f" % v for v in (0.1, 0.2, 0.3, 0.4, 0.5))
