"""Extended-XYZ reader (src/format/xyz_struc.cpp:38-186 restated in
ann_amd/xyz.py), CPU only:

* a hand-written frame in the reference's header dialect (mixed case,
  quoted lattice as columns, forces declared): atom count, names, energy,
  forces, lattice columns and positions returned to the cell (a transposed
  lattice or an unwrapped position fails);
* write -> read round trip of a multi-structure batch with energies and forces,
  and the oracle's energies of the read batch equal those of the original
  (a wrong species mapping or cell changes them);
* malformed input raises.
"""
import numpy as np
import pytest

from ann_amd import configs
from ann_amd.xyz import read_xyz, write_xyz


def test_reference_header_dialect(tmp_path):
    text = """3
Lattice="10.0 0.0 0.0 1.0 9.0 0.0 0.0 0.0 8.0" Properties=species:S:1:pos:R:3:forces:R:3 potential_energy=-12.5
Li 1.0 2.0 3.0 0.1 0.2 0.3
P 11.5 -1.0 4.0 -0.1 0.0 0.5
s 0.5 0.5 9.0 0.0 0.0 -1.0
"""
    f = tmp_path / "a.xyz"
    f.write_text(text)
    b, E, F = read_xyz(str(f), ["Li", "P", "S"])
    assert b.nstruc == 1 and b.ntotal == 3
    assert list(b.species) == [0, 1, 2]
    assert E[0] == -12.5
    assert np.allclose(F, [[0.1, 0.2, 0.3], [-0.1, 0.0, 0.5], [0.0, 0.0, -1.0]])
    R = b.cells[0]
    assert np.allclose(R[:, 0], [10, 0, 0]) and np.allclose(R[:, 1], [1, 9, 0]) and np.allclose(R[:, 2], [0, 0, 8])
    frac = np.linalg.solve(R, b.pos.T).T
    assert np.all(frac >= 0) and np.all(frac < 1)
    raw = np.array([[1.0, 2.0, 3.0], [11.5, -1.0, 4.0], [0.5, 0.5, 9.0]])
    n = np.round(np.linalg.solve(R, (raw - b.pos).T).T)
    assert np.allclose(raw - b.pos, n @ R.T)          # moved by whole lattice vectors only


def test_round_trip_and_energies(tmp_path, oracle_mod):
    model = configs.model_lips()
    batch = configs.batch_lips(2, seed=71)
    orc = oracle_mod.Oracle(model)
    E, F, _ = orc.compute_batch(batch, nthreads=4)
    f = tmp_path / "b.xyz"
    names = ["Li", "P", "S"]
    write_xyz(str(f), batch, names, energies=E, forces=F)
    b2, E2, F2 = read_xyz(str(f), names)
    assert np.array_equal(b2.natoms, batch.natoms)
    assert np.array_equal(b2.species, batch.species)
    assert np.allclose(b2.cells, batch.cells, rtol=0, atol=0)
    assert np.allclose(E2, E, rtol=0, atol=0) and np.allclose(F2, F, rtol=0, atol=0)
    Eb, _, _ = orc.compute_batch(b2, nthreads=4)
    assert np.allclose(Eb, E, rtol=1e-12, atol=1e-10)


def test_malformed_input_raises(tmp_path):
    f = tmp_path / "c.xyz"
    f.write_text("0\nLattice=\"1 0 0 0 1 0 0 0 1\"\n")
    with pytest.raises(ValueError):
        read_xyz(str(f), ["H"])
    f.write_text("1\nProperties=species:S:1:pos:R:2\nH 0 0 0\n")
    with pytest.raises(ValueError):
        read_xyz(str(f), ["H"])
    f.write_text("1\nProperties=species:S:1:pos:R:3\nX 0 0 0\n")
    with pytest.raises(ValueError):
        read_xyz(str(f), ["H"])
