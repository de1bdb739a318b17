"""The vector-ALU network kernel (ann_amd/csrc/net_vec.hip) against the FP64
oracle network on the same GPU-computed G, and against the matrix-core kernel.

Tolerances (the bars of the network parity test): atom energies
<= 1e-6 * max(1, max |E|); dE/dG relative RMS <= 1e-5; the two GPU kernels
agree to the same bars.  Shapes: widths that are not multiples of the kernel's
8-neuron chunks (20, 13, 7: padded chunks and row groups), three, two, one and
no hidden layers, three neurons, three structures of 192 atoms (partial tiles at
the ends of the species blocks), and a network too wide for the kernel's LDS
budget (64-64-64 on 80 inputs), where the matrix-core kernel runs instead.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _engine(model, batch):
    """An engine with the vector-ALU network selected (tuning entry 7 = 0; the
    default is the matrix-core kernel)."""
    from ann_amd.gpu import Engine
    eng = Engine(model, "cuda:0")
    eng.set_batch(batch)
    eng.neighbors()
    eng.set_tuning(eng.tuning()[:6] + [0])
    return eng


def _check_net(orc, species, G, E, dE):
    Gc = G.cpu().numpy().astype(np.float64)[:, :orc.ld]
    Eo, dEo = orc.net(species, Gc)
    E, dE = E.cpu().numpy(), dE.cpu().numpy()
    assert np.max(np.abs(E - Eo)) <= 1e-6 * max(1.0, np.abs(Eo).max())
    err = np.sqrt(((dE[:, :orc.ld] - dEo) ** 2).sum() / (dEo ** 2).sum())
    assert err <= 1e-5, err
    assert np.all(dE[:, orc.ld:] == 0)


@pytest.mark.parametrize("hidden,neuron", [((20, 13, 7), "SOFTPLUS"), ((32, 32), "SOFTPLUS"), ((32, 32), "TANH"),
                                           ((32, 32), "SWISH"), ((48,), "SOFTPLUS"), ((), "SOFTPLUS")])
def test_vector_network_against_oracle_and_matrix_cores(oracle_mod, hidden, neuron):
    from ann_amd import configs
    from ann_amd.gpu import FAST
    model = configs.model_water(neuron=neuron, hidden=hidden)
    batch = configs.batch_water(3, seed=41)
    print(f"hidden {hidden} neuron {neuron}", flush=True)
    eng = _engine(model, batch)
    orc = oracle_mod.Oracle(model)
    G = eng.symm(FAST)
    eng.profile(True)
    E, dE = eng.net(G, FAST)
    assert eng.profile_get("net_vec")[1] == 1
    _check_net(orc, batch.species, G, E, dE)
    # the matrix-core kernel on the same rows
    eng.set_tuning(eng.tuning()[:6] + [1])
    E2, dE2 = eng.net(G, FAST)
    assert eng.profile_get("net_mfma")[1] == 1 and eng.profile_get("net_vec")[1] == 1
    eng.profile(False)
    e1, e2 = E.cpu().numpy(), E2.cpu().numpy()
    assert np.max(np.abs(e1 - e2)) <= 1e-6 * max(1.0, np.abs(e1).max())
    d1, d2 = dE.cpu().numpy(), dE2.cpu().numpy()
    assert np.sqrt(((d1 - d2) ** 2).sum() / (d1 ** 2).sum()) <= 1e-5


def test_vector_network_energy_forces(oracle_mod):
    """The whole E+F step with the vector-ALU network (odd widths, three hidden
    layers) against the oracle's energies and forces."""
    from ann_amd import configs
    from ann_amd.gpu import FAST
    model = configs.model_water(hidden=(20, 13, 7))
    batch = configs.batch_water(3, seed=42)
    eng = _engine(model, batch)
    eng.profile(True)
    E, F, _ = eng.energy_forces(FAST)
    assert eng.profile_get("net_vec")[1] == 1
    eng.profile(False)
    Eo, Fo, _ = oracle_mod.Oracle(model).compute_batch(batch, nthreads=4)
    assert np.max(np.abs(E.cpu().numpy() - Eo) / batch.natoms) <= 1e-6
    assert np.sqrt(((F.cpu().numpy() - Fo) ** 2).sum() / (Fo ** 2).sum()) <= 1e-5


def test_wide_network_runs_on_matrix_cores(oracle_mod):
    """Weights beyond the vector kernel's 64 KiB LDS budget take the matrix-core
    kernel, with the same parity."""
    from ann_amd import configs
    from ann_amd.gpu import FAST
    model = configs.model_water(hidden=(64, 64, 64))
    batch = configs.batch_water(2, seed=43)
    eng = _engine(model, batch)
    G = eng.symm(FAST)
    eng.profile(True)
    E, dE = eng.net(G, FAST)
    assert eng.profile_get("net_vec")[1] == 0 and eng.profile_get("net_mfma")[1] == 1
    eng.profile(False)
    _check_net(oracle_mod.Oracle(model), batch.species, G, E, dE)


def test_parameter_update_reaches_vector_tables(oracle_mod):
    """New parameters (set_params) rebuild the kernel's weight tables before the
    next evaluation, also through the captured E+F graph."""
    from ann_amd import configs
    from ann_amd.gpu import FAST
    model = configs.model_water(hidden=(20, 13, 7))
    batch = configs.batch_water(2, seed=44)
    eng = _engine(model, batch)
    eng.ef_graph(True)
    for _ in range(2):
        eng.energy_forces(FAST)
    m2 = model.copy()
    rng = np.random.default_rng(45)
    p0 = m2.param_vector()
    p = p0 * (1.0 + 0.1 * rng.standard_normal(p0.shape))
    m2.set_param_vector(p)
    eng.set_params(p)
    E, F, _ = eng.energy_forces(FAST)
    E, F = E.cpu().numpy(), F.cpu().numpy()
    Eo, Fo, _ = oracle_mod.Oracle(m2).compute_batch(batch, nthreads=4)
    assert np.max(np.abs(E - Eo) / batch.natoms) <= 1e-6
    assert np.sqrt(((F - Fo) ** 2).sum() / (Fo ** 2).sum()) <= 1e-5
