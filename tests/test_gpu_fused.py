"""The fused E+F kernel (moment form: G, the atom's network and the force
adjoint on one wavefront per atom; ann_amd/csrc/symm.hip, moment_atom<3>)
against the FP64 oracle and against the separate G / network / adjoint
kernels.

Tolerances (the parity bars): energies <= 1e-6 eV/atom, forces and virial
relative RMS <= 1e-5 against the oracle; the fused and separate paths agree to
the same bars.  Cases: water (two species, zeta up to 4), Li/P/S (three species,
zeta 2), an uneven triclinic batch, tiny cells that take the fallback search,
odd network widths with three hidden layers, a network without hidden layers,
the captured E+F graph across rebuilds with moved positions, and a skin list
(entries beyond rc dropped).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

FAST = 1


def _engine(model, batch, fused, rc=None):
    from ann_amd.gpu import Engine
    eng = Engine(model, "cuda:0")
    eng.set_batch(batch)
    eng.neighbors() if rc is None else eng.neighbors(rc)
    t = eng.tuning()
    t[7] = fused
    eng.set_tuning(t)
    return eng


def _cases():
    from ann_amd import configs
    return {
        "water4": (configs.model_water(), configs.batch_water(4, seed=51)),
        "lips3": (configs.model_lips(), configs.batch_lips(3, seed=52)),
        "uneven6": (configs.model_water(), configs.batch_uneven(6, seed=53, nmax=300)),
        "tiny20": (configs.model_lips(), configs.small_random(54, natoms=20, a0=5.9)),
        "odd_widths": (configs.model_water(hidden=(20, 13, 7)), configs.batch_water(2, seed=55)),
        "no_hidden": (configs.model_water(hidden=()), configs.batch_water(2, seed=56)),
    }


@pytest.fixture(scope="module")
def cases():
    return _cases()


def _vs_oracle(oracle_mod, model, batch, E, F, V):
    Eo, Fo, Vo = oracle_mod.Oracle(model).compute_batch(batch, nthreads=4)
    E, F, V = E.cpu().numpy(), F.cpu().numpy(), V.cpu().numpy()
    assert np.max(np.abs(E - Eo) / batch.natoms) <= 1e-6
    assert np.sqrt(((F - Fo) ** 2).sum() / (Fo ** 2).sum()) <= 1e-5
    Vo = np.asarray(Vo).reshape(V.shape)
    assert np.sqrt(((V - Vo) ** 2).sum() / (Vo ** 2).sum()) <= 1e-5


@pytest.mark.parametrize("name", list(_cases().keys()))
def test_fused_against_oracle_and_separate(cases, oracle_mod, name):
    model, batch = cases[name]
    print(f"case {name}", flush=True)
    eng = _engine(model, batch, 1)
    eng.profile(True)
    E, F, V = eng.energy_forces(FAST)
    assert eng.profile_get("ef_fused")[1] == 1
    assert eng.profile_get("symm_fwd_mom")[1] == 0
    eng.profile(False)
    _vs_oracle(oracle_mod, model, batch, E, F, V)
    sep = _engine(model, batch, 0)
    E2, F2, V2 = sep.energy_forces(FAST)
    f1, f2 = F.cpu().numpy(), F2.cpu().numpy()
    assert np.max(np.abs(E.cpu().numpy() - E2.cpu().numpy()) / batch.natoms) <= 1e-6
    assert np.sqrt(((f1 - f2) ** 2).sum() / (f2 ** 2).sum()) <= 1e-5


def test_fused_graph_replays_across_rebuilds(cases):
    """The captured E+F graph with the fused kernel, replayed with positions
    moved and the list rebuilt before every call, equals eager calls bit for
    bit."""
    import torch
    model, batch = cases["water4"]
    e1, e2 = _engine(model, batch, 1), _engine(model, batch, 1)
    e2.ef_graph(True)
    dev = e1.device
    outs = [tuple(torch.empty(sh, dtype=torch.float64, device=dev)
                  for sh in ((batch.nstruc,), (batch.ntotal, 3), (batch.nstruc, 3, 3))) for _ in range(2)]
    rng = np.random.default_rng(57)
    for step in range(4):
        print(f"step {step}", flush=True)
        d = torch.tensor(rng.normal(scale=0.01, size=(batch.ntotal, 3)), device=dev)
        for e, o in ((e1, outs[0]), (e2, outs[1])):
            e.d_pos += d
            e.neighbors()
            e.energy_forces(FAST, out=o)
        torch.cuda.synchronize()
        for a, b in zip(outs[0], outs[1]):
            assert torch.equal(a, b), step
    e2.ef_graph(False)


def test_fused_with_skin_list(cases, oracle_mod):
    """A list built at rc + skin: the fused kernel drops the entries beyond rc
    (their adjoint slots zero) and matches the oracle at rc."""
    model, batch = cases["water4"]
    eng = _engine(model, batch, 1, rc=model.rc + 0.5)
    E, F, V = eng.energy_forces(FAST)
    _vs_oracle(oracle_mod, model, batch, E, F, V)


def test_fused_parameter_update(cases, oracle_mod):
    """New parameters reach the fused kernel's weight tables."""
    model, batch = cases["odd_widths"]
    eng = _engine(model, batch, 1)
    eng.energy_forces(FAST)
    m2 = model.copy()
    p0 = m2.param_vector()
    p = p0 * (1.0 + 0.1 * np.random.default_rng(58).standard_normal(p0.shape))
    m2.set_param_vector(p)
    eng.set_params(p)
    E, F, V = eng.energy_forces(FAST)
    _vs_oracle(oracle_mod, m2, batch, E, F, V)
