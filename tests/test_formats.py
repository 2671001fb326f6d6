"""On-disk formats (adscf_amd/formats.py): list-directed parsing as the drivers' READ
statements do it, the shipped input files of the three drivers, and exact write -> read
round trips of every writer."""
import os

import numpy as np
import pytest

from adscf_amd import formats as F

HERE = os.path.dirname(os.path.abspath(__file__))

# the shipped input files (Main_*/input_general, input_interactions), as data
DIBLOCK_GENERAL = """3.6d0\t\tsizex
1.d0\t\tsizey
3.6d0\t\tsizez

1.0d0\t\tC\t
2.5d0\t \tmu_scf \t! former 26.92245d0

0.001d0\tds \t\t

100000\t\titermax_scf
1.0d-12\t\taccuracy_scf

0.1d0\t\tlambda
"""
MULTIBLOCK_GENERAL = """17.2d0\t\tsizex
1.0d0\t\tsizey
1.0d0\t\tsizez

1.0d0\t\tC\t    ! Reduced in case of thin film in z direction
5.65d0 \t\tmu_scf \t

0.01d0\t\tds \t\t

0.          start_time
6\t        number_of_timesteps
0.00005\t\ttimestep

.false.\t\tqread  ! qread = .true. : read from file cnf.in
3           qinit  ! 1=cosine, 2=random densities, 3=random fields
.true.      qddft  ! qddft = .true. : DDFT simulation; otherweise SCF calculation

10.          upper omega cutoff for DDFT calculations
"""
BLEND_GENERAL = """20.0d0\t\tsizex
5.0d0\t\tsizey
1.0d0\t\tsizez

50.0d0\t\tnumber of chains A
27.0d0 \t\tmu_scf A
1.0d0\t\tlength of chain A

50.0d0\t\tnumber of chains B
27.0d0 \t\tmu_scf B
1.0d0\t\tlength of chain B

0.1d0\t\tds

10\t\tnumber of timesteps
0.001\t\ttimestep

.false.\t\tqread
3\t\tqinit
.true.\t\tqepd
"""
MULTIBLOCK_INTERACTIONS = """200.0d0        ! kappa (Helfand compressibility parameter)

70.0d0         ! chi  (Flory Huggins interaction parameter)

1.0d0          ! volume of monomer_type 1
1.0d0          ! volume of monomer_type 2
"""


def test_list_directed_values():
    assert F._split_values("1.0d0, 2*3.5 .t. / 9") == ["1.0d0", "3.5", "3.5", ".t."]
    assert F.parse_real("2.5D-3") == 2.5e-3 and F.parse_real("5.0000000745058060E-002") == 5.0000000745058060e-2
    assert F.parse_real("3.") == 3.0 and F.parse_real("-7") == -7.0
    assert F.parse_logical(".true.") and F.parse_logical("T") and not F.parse_logical(".f")
    with pytest.raises(ValueError):
        F.parse_logical("x")


def test_record_reader_semantics(tmp_path):
    p = tmp_path / "r"
    p.write_text("1 2 3\n\n4\n5 6\n")
    r = F.RecordReader(str(p))
    assert r.read(int) == 1          # the rest of the record is dropped
    r.read()                         # skips one record
    assert r.read(int, int) == (4, 5)  # a READ continues onto the next record
    with pytest.raises(EOFError):
        r.read(int)


def test_shipped_inputs(tmp_path):
    (tmp_path / "d").write_text(DIBLOCK_GENERAL)
    d = F.read_input_general(str(tmp_path / "d"), "diblock")
    assert (d.sizex, d.sizey, d.sizez, d.C, d.mu, d.ds) == (3.6, 1.0, 3.6, 1.0, 2.5, 0.001)
    assert d.itermax_scf == 100000 and d.accuracy_scf == 1e-12 and d.lam == 0.1
    (tmp_path / "m").write_text(MULTIBLOCK_GENERAL)
    m = F.read_input_general(str(tmp_path / "m"), "multiblock")
    assert (m.sizex, m.C, m.mu, m.ds, m.start_time) == (17.2, 1.0, 5.65, 0.01, 0.0)
    assert m.number_of_timesteps == 6 and m.timestep == 5e-5
    assert (m.qread, m.qinit, m.qddft, m.wcutoff) == (False, 3, True, 10.0)
    (tmp_path / "b").write_text(BLEND_GENERAL)
    b = F.read_input_general(str(tmp_path / "b"), "blend")
    assert b.number_of_chains == [50.0, 50.0] and b.mu == [27.0, 27.0] and b.chain_length == [1.0, 1.0]
    assert (b.ds, b.number_of_timesteps, b.timestep, b.qread, b.qinit, b.qepd) == (0.1, 10, 0.001, False, 3, True)
    (tmp_path / "i").write_text(MULTIBLOCK_INTERACTIONS)
    it = F.read_interactions(str(tmp_path / "i"))
    assert it.kappa == 200.0 and it.chi[0, 1] == 70.0 and it.chi[1, 0] == 70.0 and it.volume == [1.0, 1.0]


def test_shipped_mobilities():
    m = F.read_mobilities(os.path.join(HERE, "golden", "multiblock_input_mobilities"))
    assert m.nq == 200 and m.dq == 0.05
    assert m.table[0, 0, 0] == 0.56173199034771881
    assert m.table[1, 0, 0] == 0.13768489198101524 and m.table[0, 1, 0] == 0.13768489198101524
    assert m.table[1, 1, 199] == 8.5418864569320838e-2


def test_mobilities_checks(tmp_path):
    m = F.Mobilities(3, 0.1, np.arange(12, dtype=float).reshape(2, 2, 3))
    m.table[0, 1] = m.table[1, 0]
    F.write_mobilities(str(tmp_path / "mob"), m)
    back = F.read_mobilities(str(tmp_path / "mob"))
    assert back.nq == 3 and back.dq == 0.1 and np.array_equal(back.table, m.table)
    bad = (tmp_path / "mob").read_text().replace("2.0000000000000001E-01", "2.5E-01")
    (tmp_path / "bad").write_text(bad)
    with pytest.raises(ValueError, match="not consistent"):
        F.read_mobilities(str(tmp_path / "bad"))
    (tmp_path / "big").write_text("1001\n0.1\n")
    with pytest.raises(ValueError, match="nqmax"):
        F.read_mobilities(str(tmp_path / "big"))


def test_round_trips(tmp_path):
    rng = np.random.default_rng(0)
    for inp in (F.DiblockInput(3.6, 1.0, 3.6, 1.0, 2.5, 0.001, 100000, 1e-12, 0.1),
                F.BlendInput(20.0, 5.0, 1.0, [50.0, 40.0], [27.0, 26.5], [1.0, 0.5], 0.1, 10, 0.001, True, 2, False),
                F.MultiblockInput(17.2, 1.0, 1.0, 1.0, 5.65, 0.01, 0.25, 6, 5e-5, False, 3, True, 10.0)):
        F.write_input_general(str(tmp_path / "g"), inp)
        kind = {F.DiblockInput: "diblock", F.BlendInput: "blend", F.MultiblockInput: "multiblock"}[type(inp)]
        assert F.read_input_general(str(tmp_path / "g"), kind) == inp
    inter = F.Interactions(25.0, np.array([[0.0, 1.5, 2.5], [1.5, 0.0, 3.5], [2.5, 3.5, 0.0]]), [1.0, 0.9, 1.1])
    F.write_interactions(str(tmp_path / "i"), inter)
    back = F.read_interactions(str(tmp_path / "i"), components=3)
    assert back.kappa == 25.0 and np.array_equal(back.chi, inter.chi) and back.volume == inter.volume
    field = rng.standard_normal((2, 5, 3, 4)) * 10.0 ** rng.integers(-20, 20, (2, 5, 3, 4))
    F.write_cnf(str(tmp_path / "cnf"), field)
    assert np.array_equal(F.read_cnf(str(tmp_path / "cnf"), field.shape), field)
    coords = (np.arange(1, 5) * 0.5, np.arange(1, 4) * 1.0, np.arange(1, 6) * 0.25)
    for order in ("zyx", "xyz"):
        F.write_profile(str(tmp_path / "p"), field, coords, order, header=["# Energy :  -1.0", ""])
        assert np.array_equal(F.read_profile(str(tmp_path / "p"), field.shape, order), field)
    F.write_profile(str(tmp_path / "pt"), field, coords, "xyz", leading=[0.1])
    assert np.array_equal(F.read_profile(str(tmp_path / "pt"), field.shape, "xyz", index_columns=4), field)


def test_cnf_reads_repeat_counts_and_fortran_exponents(tmp_path):
    (tmp_path / "c").write_text(" 3*1.5D0, 2.0d-1\n -4.0E+000 /\n")
    a = F.read_cnf(str(tmp_path / "c"), (1, 1, 1, 5))
    assert np.array_equal(a.ravel(), [1.5, 1.5, 1.5, 0.2, -4.0])


def test_profile_with_wrapped_records(tmp_path):
    """flang continues a long list-directed WRITE on the next record; blank records end each
    sweep and a '#' header precedes the rows (final_dens.dat, main_multiblock.f90:345-367)."""
    vals = np.array([[[[0.1, 0.2]]], [[[0.3, 0.4]]]])  # (2, 1, 1, 2)
    text = (" # Energy :   -8.5\n\n"
            " 0.20000000000000001        1.0000000000000000        1.0000000000000000\n"
            "   0.10000000000000001        0.29999999999999999\n\n"
            " 0.40000000000000002  1.0 1.0 0.2 0.4\n\n")
    (tmp_path / "f").write_text(text)
    got = F.read_profile(str(tmp_path / "f"), vals.shape, "xyz")
    assert np.array_equal(got[:, 0, 0, :], [[0.1, 0.2], [0.29999999999999999, 0.4]])
