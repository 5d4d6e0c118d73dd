"""XYZ input through the C ABI (gfn1_system_from_xyz), the reader of
PeriodicSystem::from_xyz_str (src/system.rs:45-99, symbol_to_z :154-165).
Host-only parsing: runs without a GPU."""
import numpy as np
import pytest

from gfn1_amd import NativeError, system_from_xyz, system_from_xyz_file

A2B = 1.8897261246257702  # system.rs:8


def test_reference_h2_text():
    # the stretched-H2 input of tests/optimizer.rs:9-13
    z, x, q = system_from_xyz("2\nH2 stretched\nH 0.000000 0.000000 0.000000\nH 1.100000 0.000000 0.000000\n", 0.0)
    assert list(z) == [1, 1] and q == 0.0
    assert np.array_equal(x, np.array([[0.0, 0.0, 0.0], [1.1 * A2B, 0.0, 0.0]]))


def test_units_symbols_and_line_endings(tmp_path):
    text = "3\r\nwater\r\no 0 0 0\r\nh 0.757 0.586 0\r\n1 -0.757 0.586 0\r\n"
    z, x, q = system_from_xyz(text, -1.0, coordinates_are_bohr=True)
    assert list(z) == [8, 1, 1] and q == -1.0
    assert np.array_equal(x[1], [0.757, 0.586, 0.0])
    z2, x2, _ = system_from_xyz("2\n\nHE 0 0 0\nCl 1 2 3 extra fields\n")
    assert list(z2) == [2, 17] and np.allclose(x2[1], np.array([1.0, 2.0, 3.0]) * A2B, rtol=0, atol=0)
    p = tmp_path / "m.xyz"
    p.write_text("1\ncomment\nC 0.5 0 0\n")
    zf, xf, _ = system_from_xyz_file(str(p))
    assert list(zf) == [6] and xf[0, 0] == 0.5 * A2B


@pytest.mark.parametrize("text,kind,fragment", [
    ("", "InvalidInput", "empty XYZ"),
    ("two\nc\n", "InvalidInput", "invalid XYZ atom count"),
    ("2\nc\nH 0 0 0\n", "InvalidInput", "ended before atom 1"),
    ("1\nc\nH 0 0\n", "InvalidInput", "line 3 has fewer than 4 fields"),
    ("1\nc\nXx 0 0 0\n", "InvalidInput", "unknown element on line 3: Xx"),
    ("1\nc\n87 0 0 0\n", "InvalidInput", "unknown element"),
    ("1\nc\nH 0 zero 0\n", "Parse", "invalid floating point value: zero"),
    ('1\nLattice="3 0 0 0 3 0 0\nH 0 0 0\n', "InvalidInput", "unterminated Lattice"),
    ('1\nLattice="3 0 0 0 3 0 0 0"\nH 0 0 0\n', "InvalidInput", "needs 9 numbers, got 8"),
    ('1\nLattice="3 0 0 0 3 0 0 0 x"\nH 0 0 0\n', "InvalidInput", "invalid number in XYZ Lattice"),
    ('1\nLattice="3 0 0 0 3 0 0 0 3" pbc="T T"\nH 0 0 0\n', "InvalidInput", "needs 3 boolean tokens"),
    ('1\nLattice="3 0 0 0 3 0 0 0 3" pbc="T T maybe"\nH 0 0 0\n', "InvalidInput", "invalid XYZ pbc token"),
    ('1\nLattice="3 0 0 0 3 0 0 0 3" pbc="T T F"\nH 0 0 0\n', "InvalidInput", "fully periodic"),
    ('1\nLattice="3 0 0 0 3 0 6 0 0"\nH 0 0 0\n', "SingularCell", "singular"),
])
def test_errors(text, kind, fragment):
    with pytest.raises(NativeError) as e:
        system_from_xyz(text)
    assert e.value.kind == kind
    assert fragment in str(e.value)


def test_extended_xyz_lattice():
    # system.rs:191-219: Lattice="a b c" in the coordinate unit, pbc optional
    text = '2\nLattice="5.6 0 0 0 5.6 0 0.1 0.2 5.5" pbc="T T T" Properties=species:S:1:pos:R:3\nNa 0 0 0\nCl 2.8 2.8 2.8\n'
    z, x, q, cell = system_from_xyz(text, with_lattice=True)
    assert list(z) == [11, 17]
    assert np.array_equal(cell, np.array([[5.6, 0, 0], [0, 5.6, 0], [0.1, 0.2, 5.5]]) * A2B)
    z, x, q, cell = system_from_xyz(text.replace(' pbc="T T T"', ''), coordinates_are_bohr=True, with_lattice=True)
    assert cell[2, 2] == 5.5
    assert system_from_xyz("1\nc\nH 0 0 0\n", with_lattice=True)[3] is None
