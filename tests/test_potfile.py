"""The NNP potential text format (NNP::write / NNP::read, src/nnp/nnp.cpp:197-307):
Python writer/reader round trip, and the C++ drop-in reader/writer producing
the same text (no GPU needed for these)."""
import os
import subprocess

import numpy as np
import pytest

from ann_amd import configs, potfile

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "tests", "cpp", "test_nnp_gpu")


@pytest.mark.parametrize("make", [configs.model_ar256, configs.model_water, configs.model_lips])
def test_python_roundtrip(tmp_path, make):
    m = make()
    p = tmp_path / "pot.nnp"
    potfile.write(str(p), m)
    r = potfile.read(str(p))
    assert r.species == m.species and r.rc == m.rc and r.cutoff == m.cutoff
    assert np.allclose(r.evac, m.evac, atol=1e-6)
    for a in range(m.nspecies):
        for b in range(m.nspecies):
            assert r.radial[a][b][0] == m.radial[a][b][0]
            assert np.allclose(r.radial[a][b][1], m.radial[a][b][1], atol=1e-15)
        for q in range(len(m.angular[a])):
            for f in range(4):
                assert np.allclose(r.angular[a][q][f], m.angular[a][q][f], atol=1e-15)
        assert np.allclose(r.nets[a].params, m.nets[a].params, atol=1e-15)
        assert r.nets[a].nodes == m.nets[a].nodes and r.nets[a].neuron == m.nets[a].neuron


def test_file_order_differs_from_storage_order_for_three_species(tmp_path):
    """Angular blocks are written (0,0),(0,1),(0,2),(1,1),... but stored in
    TensorTri order (0,0),(0,1),(1,1),(0,2),...: give every pair a distinct
    basis and check each lands on its own pair after the round trip."""
    m = configs.model_lips()
    for a in range(3):
        for p in range(6):
            name, eta, zeta, lam = m.angular[a][p]
            m.angular[a][p] = (name, eta + 0.01 * (p + 1), zeta, lam)
    path = tmp_path / "pot.nnp"
    potfile.write(str(path), m)
    r = potfile.read(str(path))
    for p in range(6):
        assert np.allclose(r.angular[0][p][1], m.angular[0][p][1])


@pytest.mark.skipif(not os.path.exists(DRIVER), reason="C++ drop-in driver not built")
def test_cpp_reader_writer_matches_python(tmp_path):
    m = configs.model_lips()
    a, b = tmp_path / "a.nnp", tmp_path / "b.nnp"
    potfile.write(str(a), m)
    out = subprocess.run([DRIVER, "roundtrip", str(a), str(b)], capture_output=True, text=True, timeout=60)
    assert out.returncode == 0, out.stderr
    assert open(a).read() == open(b).read()
    assert "params 21171" in out.stdout
