"""MP_TOOLS .dat reader and the host side of the mp_sql path (CPU).

The snapshot tests/golden/al_u/Al_U_large_n0001.dat is the reference's own
example input (examples/amorphous_Al92U08/data); its header and species
partition are known answers: 27,600 Al + 2,400 U, box 79.8159637 A, identity
cell matrices, at_ind(1) = serial number, at_ind(4) = species
(code/docs/1_mp_data_doc_154.txt:65-113).
"""
import os

import numpy as np
import pytest

from mp_tools_amd import mpdat, sql

HERE = os.path.dirname(os.path.abspath(__file__))
DAT = os.path.join(HERE, "golden", "al_u", "Al_U_large_n0001.dat")


def test_namelist_parser():
    d = mpdat.parse_namelist('&H  A= 3*1.5 , N=2, S="ab c  ", V= 1.0, 2.5E-01, /')
    assert d["A"] == [1.5, 1.5, 1.5] and d["N"] == 2 and d["S"] == "ab c  " and d["V"] == [1.0, 0.25]


def test_read_al_u_snapshot():
    s = mpdat.read_dat(DAT)
    assert s.version == "1.56" and s.n_head == 4 and s.n_tot == 30000
    assert list(s.species_counts) == [27600, 2400] and s.species_names == ["Al", "U"]
    np.testing.assert_allclose(s.occupancies, [0.92, 0.08], rtol=1e-6)
    np.testing.assert_allclose(s.a_par, [79.8159637] * 3, rtol=1e-7)
    assert np.array_equal(s.a_cell_inv, np.eye(3))
    assert np.array_equal(s.at_ind[:, 0], np.arange(1, 30001))
    assert np.array_equal(s.at_ind[:, 3], np.repeat([1, 2], [27600, 2400]))
    half = s.a_par[0] / 2
    assert np.all(np.abs(s.at_pos[:, :3]) <= half)


def test_sql_geometry_of_the_reference_run():
    """Q range 20 1/A on the 79.8 A box: 509^3 modes and 2001 bins, as the
    reference's log reports ("NUFFT_3D requires about 8 Gb")."""
    s = mpdat.read_dat(DAT)
    g = sql.geometry(s.a_par, s.a_cell_inv, s.header["N_ROW"], s.n_tot, 20.0, 0.01)
    assert g.n_qq == (509, 509, 509) and g.n_hist == 2001
    np.testing.assert_allclose(float(g.dq[0]), 2 * np.pi / 79.8159637, rtol=1e-7)
    np.testing.assert_allclose(g.at_volume, 79.8159637 ** 3 / 30000, rtol=1e-6)


def test_xray_ff_closed_forms():
    ff = np.load(os.path.join(HERE, "golden", "al_u", "al_u_sql_golden.npz"))["ffpar"]
    f = sql.xray_ff(ff, np.array([0.0, 4 * np.pi]))
    np.testing.assert_allclose(f[:, 0], ff[:, [0, 2, 4, 6, 8]].sum(axis=1))        # f(0) = sum a + c
    s2 = 1.0                                                                       # (Q/4pi)^2 at Q = 4 pi
    np.testing.assert_allclose(f[:, 1], ff[:, 8] + sum(ff[:, 2 * k] * np.exp(-ff[:, 2 * k + 1] * s2) for k in range(4)))


def test_sql_output_faber_ziman_of_an_ideal_gas():
    """Uncorrelated atoms: every partial histogram equals its asymptote
    (diagonal occ_i, off-diagonal 0) at every Q, so Z(Q) and F(Q) are 0 for
    every pair and the total S(Q) is 1, whatever the weights and smoothing."""
    g = sql.geometry([50.0] * 3, np.eye(3), [1, 1, 1], 1000, 5.0, 0.01)
    occ = np.array([0.7, 0.3])
    h = np.zeros((2, 2, g.n_hist))
    h[0, 0, :], h[1, 1, :] = occ[0], occ[1]
    ff = np.load(os.path.join(HERE, "golden", "al_u", "al_u_sql_golden.npz"))["ffpar"]
    q, tot, z = sql.output(h, g, 5, occ, "xray", ffpar=ff, smooth_fwhm=5)
    assert tot is None and np.abs(z).max() < 1e-12
    q, tot, s = sql.output(h, g, 4, occ, "xray", ffpar=ff, smooth_fwhm=5)
    assert np.abs(tot - 1).max() < 1e-12
    q, tot, f = sql.output(h, g, 3, occ, "xray", ffpar=ff, smooth_fwhm=5)
    assert np.abs(tot).max() < 1e-12 and np.abs(f).max() < 1e-12
    q, tot, s = sql.output(h, g, 4, occ, "neutron", b_coh=[3.4, 8.4], smooth_fwhm=1)
    # without smoothing the first written row is one the reference zeroed
    # before weighting (mp_sql56.f90:1404 vs i_start at :1814), so it holds
    # only the asymptote terms
    assert np.abs(tot[1:] - 1).max() < 1e-12
    assert abs(q[0] - 0.49) < 1e-6                # first written row: bin int(.5/q_step) + n_smooth/2 (1-based)


def _al_u_header_kw(text_header):
    """The run's parameters as the reference's Al/U tables print them; the
    time stamp is that table's own."""
    stamp = " ".join(text_header[0].split()[-2:])
    return dict(subst_name="Al_U", mp_tool="MP_SQL 1.56", time_stamp=stamp,
                file_dat_t0="./data/Al_U_large_n0001.dat", file_dat="./data/Al_U_large_n0001.dat",
                nfile_step=1, proc_num=2, n_row=(1, 1, 1), a_par=(np.float32("79.8159637"),) * 3,
                at_names=("Al", "U"), occ=(0.92, 0.08), q_step=0.01, smooth_fwhm=5)


def _check_header(got, want):
    # record 1: the example's binary spaced the substance record differently
    # from the source's (shells(1) is written untrimmed there); compare its words
    assert got[0].split() == want[0].split()
    for k in range(1, len(want)):
        assert got[k] == want[k], (k, got[k], want[k])


@pytest.mark.parametrize("key,mode", [("Z", 5), ("I", 6), ("S", 4), ("Iu", 7)])
def test_sql_text_records_of_the_reference_tables(key, mode):
    """mp_sql's text file (mp_sql56.f90:1800-1829): the header is rebuilt from
    the run's parameters (the X-ray amplitudes from the form factors), and the
    reference's own data records, parsed back to REAL(4), are rewritten
    unchanged."""
    import json
    gd = os.path.join(HERE, "golden", "al_u")
    text = json.load(open(os.path.join(gd, "al_u_sql_text.json"), encoding="utf-8"))[key]
    gold = np.load(os.path.join(gd, "al_u_sql_golden.npz"))
    kw = _al_u_header_kw(text["header"])
    amp = sql.header_amplitudes("xray", kw["occ"], ffpar=gold["ffpar"])
    assert ["%8.4f" % a for a in amp] == [" 19.3117", " 25.1244"]
    ref = gold[key]
    assert ref.shape[0] == text["n_rows"]
    for sl, rows in ((slice(0, 3), text["first_rows"]), (slice(-3, None), text["last_rows"])):
        r = ref[sl]
        tot = None if mode == 5 else r[:, 1]
        parts = r[:, 1:] if mode == 5 else r[:, 2:]
        out = sql.table_text(r[:, 0], tot, parts, mode, amplitudes=amp, **kw).split("\n")
        hdr = text["header"]
        _check_header(out[:len(hdr)], hdr)
        assert out[len(hdr):len(hdr) + 3] == rows


def test_dat_writer_reproduces_the_reference_snapshot():
    """The .dat writer (mp_lbin56.f90:1436-1480, gfortran namelist records)
    rewrites the reference's own Al/U snapshot byte for byte."""
    p = os.path.join(HERE, "golden", "al_u", "Al_U_large_n0001.dat")
    s = mpdat.read_dat(p)
    assert s.time_stamp == "2023/11/28 10:18:08"
    assert [g for g, _ in s.groups] == ["DATA_HEADER_1", "DATA_HEADER_2", "DATA_HEADER_3"]
    assert mpdat.dat_bytes(s) == open(p, "rb").read()


def test_namelist_record_closed_forms():
    rec = mpdat.namelist_record("grp", {"A": [1.0, 1.0, 2.0], "N": 7, "S": ["ab", "ab"], "C": "x "})
    assert rec == '&GRP  A= 2*1.00000000      ,  2.00000000    ,  N=7          ,  S= 2*"ab",  C="x ",  /'
    # five values of one object, then two more blanks
    rec = mpdat.namelist_record("g", {"V": [1, 2, 3, 4, 5, 6]})
    assert rec == "&G  V=" + ",".join(("%d" % v).ljust(11) for v in range(1, 6)) + ",  " + "6".ljust(11) + ",  /"
    g, d = mpdat.parse_namelist_group(rec)
    assert g == "G" and d["V"] == [1, 2, 3, 4, 5, 6]


def test_dat_round_trip_and_trajectory_loading(tmp_path):
    """write -> read of a snapshot whose arrays end inside a record, and the
    frame stack of the trajectory loader (species-major float4 per frame)."""
    s = mpdat.read_dat(os.path.join(HERE, "golden", "al_u", "Al_U_large_n0001.dat"))
    n = 1000  # 4000 words: the last record is partly padding
    rng = np.random.default_rng(3)
    s2 = mpdat.MPSnapshot(version=s.version, n_head=s.n_head, time_stamp="2026/01/01 00:00:00",
                          groups=[(g, dict(d)) for g, d in s.groups],
                          at_ind=s.at_ind[:n].copy(), at_pos=rng.random((n, 4), dtype=np.float32))
    d1 = dict(s2.groups[0][1])
    d1.update(N_TOT=n, N_REC=-(-4 * n // mpdat.L_REC))
    d2 = dict(s2.groups[1][1])
    d2.update(NSUPER_R=[900, 100])
    s2.groups = [("DATA_HEADER_1", d1), ("DATA_HEADER_2", d2), s2.groups[2]]
    paths = []
    for k in range(3):
        p = str(tmp_path / f"f{k}.dat")
        s2.at_pos = (rng.random((n, 4)) * (k + 1)).astype(np.float32)
        mpdat.write_dat(p, s2)
        back = mpdat.read_dat(p)
        assert back.n_tot == n and list(back.species_counts) == [900, 100]
        np.testing.assert_array_equal(back.at_pos, s2.at_pos)
        np.testing.assert_array_equal(back.at_ind, s2.at_ind)
        assert os.path.getsize(p) == (4 + 2 * d1["N_REC"]) * 4 * mpdat.L_REC
        paths.append((p, s2.at_pos.copy()))
    traj, first = mpdat.load_trajectory([p for p, _ in paths], device="cpu")
    assert tuple(traj.shape) == (3, n, 4)
    for k, (_, pos) in enumerate(paths):
        np.testing.assert_array_equal(traj[k].numpy(), pos)
    other = str(tmp_path / "other.dat")
    mpdat.write_dat(other, s)
    with pytest.raises(ValueError):
        mpdat.load_trajectory([paths[0][0], other], device="cpu")


@pytest.mark.parametrize("n_traj,shells", [(1, 0), (2, 0), (1, 1), (2, 1), (0, 1)])
def test_dat_velocity_force_shell_blocks(tmp_path, n_traj, shells):
    """The blocks after AT_POS in the writer's order (mp_lbin56.f90:1490-1547)
    round-trip, and sit where mp_dos reads the velocities: the cores' right
    after AT_POS, the shells' (n_traj + 1) blocks after it
    (mp_dos56.f90:429-452)."""
    s = mpdat.read_dat(os.path.join(HERE, "golden", "al_u", "Al_U_large_n0001.dat"))
    n = 700
    rng = np.random.default_rng(n_traj * 10 + shells)
    d1 = dict(s.groups[0][1])
    d1.update(N_TOT=n, N_REC=-(-4 * n // mpdat.L_REC), N_TRAJ=n_traj, J_SHELL_OUT=shells)
    d2 = dict(s.groups[1][1])
    d2.update(NSUPER_R=[600, 100])
    blk = lambda: rng.random((n, 4), dtype=np.float32)
    s2 = mpdat.MPSnapshot(version=s.version, n_head=s.n_head, time_stamp="2026/01/01 00:00:00",
                          groups=[("DATA_HEADER_1", d1), ("DATA_HEADER_2", d2), s.groups[2]],
                          at_ind=s.at_ind[:n].copy(), at_pos=blk())
    names = (["at_vel"] if n_traj >= 1 else []) + (["at_force"] if n_traj == 2 else [])
    if shells:
        names += ["at_pos_s"] + (["at_vel_s"] if n_traj >= 1 else []) + (["at_force_s"] if n_traj == 2 else [])
    for nm in names:
        setattr(s2, nm, blk())
    p = str(tmp_path / "b.dat")
    mpdat.write_dat(p, s2)
    n_rec = d1["N_REC"]
    assert os.path.getsize(p) == (4 + (2 + len(names)) * n_rec) * 4 * mpdat.L_REC
    back = mpdat.read_dat(p)
    for nm in ["at_pos", "at_vel", "at_force", "at_pos_s", "at_vel_s", "at_force_s"]:
        want = getattr(s2, nm)
        got = getattr(back, nm)
        assert (want is None and got is None) or np.array_equal(got, want), nm
    raw = open(p, "rb").read()
    rb = 4 * mpdat.L_REC
    if n_traj >= 1:
        # mp_dos: cores' velocities right after AT_POS; shells' after (n_traj + 1) more blocks
        o = (4 + 2 * n_rec) * rb
        assert np.array_equal(np.frombuffer(raw, np.float32, 4 * n, o).reshape(n, 4), back.at_vel)
        if shells:
            o = (4 + 2 * n_rec + (n_traj + 1) * n_rec) * rb
            assert np.array_equal(np.frombuffer(raw, np.float32, 4 * n, o).reshape(n, 4), back.at_vel_s)
        traj, _ = mpdat.load_trajectory([p, p], device="cpu", block="at_vel")
        assert np.array_equal(traj[1].numpy(), s2.at_vel)
    else:
        with pytest.raises(ValueError):
            mpdat.load_trajectory([p], device="cpu", block="at_vel")
    s2.at_pos_s = None if shells else s2.at_pos_s
    if shells:
        with pytest.raises(ValueError):
            mpdat.dat_bytes(s2)
