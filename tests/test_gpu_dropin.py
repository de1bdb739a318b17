"""The C++ drop-in classes (ann_amd/cpp, the reference's NNP / NeighborList /
Structure interface over the C ABI) on the GPU: the reference force test
restated in C++ (tests/cpp/test_nnp_gpu.cpp): forces from NNP::compute
against central differences of NNP::symm + NNP::energy (FP32 energies,
eps 1e-3 A, tolerance 2e-3 max|F|), NNP::force = NNP::compute, energy =
compute + vacuum energies, sum F = 0."""
import os
import subprocess

import pytest

from ann_amd import configs, potfile

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "tests", "cpp", "test_nnp_gpu")


@pytest.mark.parametrize("form", ["fast", "plain"])
@pytest.mark.parametrize("make,seed", [(configs.model_lips, 1), (configs.model_water, 2)])
def test_cpp_dropin_forces(tmp_path, make, seed, form):
    m = make()
    m.evac = m.evac + 0.25 * (1 + len(m.species))    # non-zero vacuum energies
    pot = tmp_path / "pot.nnp"
    potfile.write(str(pot), m)
    env = dict(os.environ, ANN_FORM=form)
    print(f"driver {make.__name__} seed {seed} form {form}", flush=True)
    out = subprocess.run([DRIVER, "force", str(pot), str(seed)], capture_output=True, text=True, timeout=240, env=env)
    print(out.stdout, out.stderr, flush=True)
    assert out.returncode == 0
