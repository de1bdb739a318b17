"""Full-size correctness of the 1,048,576-atom MD workload: with zero jitter
the 16 x 16 x 16 replica of the 256-atom water cell is the cell itself
repeated, so its energy is 4096 times the cell's and every copy of an atom
carries the cell atom's force.  The cell is evaluated by the FP64 oracle.

Bars: energy <= 1e-6 eV/atom; forces <= 1e-5 relative RMS over all copies;
virial <= 1e-4 relative.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_million_atom_replica_against_the_cell_oracle(oracle_mod):
    from ann_amd import configs
    from ann_amd.gpu import Engine, FAST
    model = configs.model_water()
    cell = configs.struc_water_md(reps=1, jitter=0.0)
    big = configs.struc_water_md(reps=16, jitter=0.0)
    assert big.ntotal == 4096 * cell.ntotal == 1048576
    Eo, Fo, Vo = oracle_mod.Oracle(model).compute_batch(cell, nthreads=4)
    print("1M-atom replica: build + E/F", flush=True)
    eng = Engine(model, "cuda:0")
    eng.set_batch(big)
    eng.neighbors()
    E, F, V = eng.energy_forces(FAST)
    E, F, V = E.cpu().numpy(), F.cpu().numpy(), V.cpu().numpy()
    de = abs(E[0] - 4096 * Eo[0]) / big.ntotal
    # replica copy c of cell atom a sits at c * 256 + a
    Fc = F.reshape(4096, cell.ntotal, 3)
    rf = float(np.sqrt(((Fc - Fo[None]) ** 2).sum() / (4096 * (Fo ** 2).sum())))
    rv = float(np.abs(V[0] - 4096 * Vo[0]).max() / (4096 * np.abs(Vo[0]).max()))
    print(f"  dE/atom={de:.2e} force relRMS={rf:.2e} virial rel={rv:.2e}", flush=True)
    assert de <= 1e-6 and rf <= 1e-5 and rv <= 1e-4
