"""Networks wider than 64 neurons (up to the 128 the kernels support) against
the FP64 oracle: E+F in all three forms, the network's energies and dE/dG on
the same G, and a full training step (loss terms and parameter gradient,
second backward included).

Tolerances (the parity bars): energies <= 1e-6 eV/atom and forces <= 1e-5
relative RMS; network energies <= 1e-6 max(1, |E|) and dE/dG <= 1e-5 relative
RMS; loss terms within the change those energy and force bars allow (the
bound is written in the test) and gradient relative (L2) <= 1e-5.  The
vector-ALU and fused paths stop at 64 neurons, so these shapes run on the
matrix-core kernels (and the PLAIN kernels).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _engine(model, batch):
    from ann_amd.gpu import Engine
    eng = Engine(model, "cuda:0")
    eng.set_batch(batch)
    eng.neighbors()
    return eng


def _water_wide():
    from ann_amd import configs
    return configs.model_water(hidden=(128, 96)), configs.batch_water(2, seed=61)


def _lips_wide():
    from ann_amd import configs
    from ann_amd.model import init_network
    m = configs.model_lips()
    rng = np.random.default_rng(62)
    m.nets = [init_network(96, [128, 72, 100], n.neuron, rng) for n in m.nets]
    return m, configs.batch_lips(2, seed=63)


@pytest.mark.parametrize("form", [0, 1, 2])
def test_wide_energy_forces(oracle_mod, form):
    model, batch = _water_wide()
    print(f"form {form}", flush=True)
    eng = _engine(model, batch)
    E, F, _ = eng.energy_forces(form)
    Eo, Fo, _ = oracle_mod.Oracle(model).compute_batch(batch, nthreads=4)
    assert np.max(np.abs(E.cpu().numpy() - Eo) / batch.natoms) <= 1e-6
    assert np.sqrt(((F.cpu().numpy() - Fo) ** 2).sum() / (Fo ** 2).sum()) <= 1e-5


@pytest.mark.parametrize("form", [0, 1])
def test_wide_network_parity(oracle_mod, form):
    model, batch = _lips_wide()
    eng = _engine(model, batch)
    G = eng.symm(form)
    E, dE = eng.net(G, form)
    orc = oracle_mod.Oracle(model)
    Gc = G.cpu().numpy().astype(np.float64)[:, :orc.ld]
    Eo, dEo = orc.net(batch.species, Gc)
    E, dE = E.cpu().numpy(), dE.cpu().numpy()
    assert np.max(np.abs(E - Eo)) <= 1e-6 * max(1.0, np.abs(Eo).max())
    err = np.sqrt(((dE[:, :orc.ld] - dEo) ** 2).sum() / (dEo ** 2).sum())
    assert err <= 1e-5, err


def test_wide_second_backward(oracle_mod):
    import torch
    model, batch = _lips_wide()
    eng = _engine(model, batch)
    G = eng.symm(1)
    rng = np.random.default_rng(65)
    v = rng.normal(size=(batch.ntotal, eng.ldg)).astype(np.float32) * 0.1
    gE = rng.normal(size=batch.nstruc)
    g = eng.net_bwd2(G, torch.tensor(v, device="cuda:0"), torch.tensor(gE, device="cuda:0")).cpu().numpy()
    orc = oracle_mod.Oracle(model)
    Gc = G.cpu().numpy().astype(np.float64)
    poff = np.concatenate([[0], np.cumsum([orc.nparams(a) for a in range(model.nspecies)])])
    ref = np.zeros(poff[-1])
    sof = np.repeat(np.arange(batch.nstruc), batch.natoms)
    for i in range(batch.ntotal):
        a = batch.species[i]
        ref[poff[a]:poff[a + 1]] += orc.nn_grad_rop(a, Gc[i, :orc.ld], v[i, :orc.ld].astype(np.float64), gE[sof[i]])
    err = np.linalg.norm(g - ref) / np.linalg.norm(ref)
    assert err <= 1e-5, err


@pytest.mark.parametrize("form", [0, 1, 2])
def test_wide_train_step(oracle_mod, form):
    import torch
    model, batch = _lips_wide()
    print(f"form {form}", flush=True)
    eng = _engine(model, batch)
    rng = np.random.default_rng(64)
    orc = oracle_mod.Oracle(model)
    E0, F0, _ = orc.compute_batch(batch, nthreads=4)
    evac = np.array([sum(model.evac[s] for s in batch.struc(k)[2]) for k in range(batch.nstruc)])
    Eref = E0 - evac + rng.normal(size=batch.nstruc) * 0.05 * batch.natoms
    Fref = F0 + rng.normal(size=F0.shape) * 0.05
    loss, g = eng.train_step(torch.tensor(Eref, device="cuda:0"), torch.tensor(Fref, device="cuda:0"),
                             1.0, 0.1, 0.0, form)
    loss, g = loss.cpu().numpy(), g.cpu().numpy()
    wE, wF = 1.0, 0.1
    L, gr, _, _, parts = orc.train_loss_grad(batch, Eref, Fref, wE, wF)
    # loss terms to the bound the energy and force bars imply: with per-atom
    # energy errors d <= 1e-6 and force errors e of total square at most
    # (1e-5)^2 sum F^2, a term sum_x w_x r_x^2 moves by at most
    # 2 sqrt(L) sqrt(D) + D, D = sum_x w_x err_x^2 (Cauchy-Schwarz); here
    # D_E = wE 1e-12 and D_F <= max_x w_x 1e-10 sum F^2, w_x = wF / (3 N_k B)
    n = np.diff(batch.atom_off)
    DE = wE * 1e-12
    DF = wF / (3.0 * n.min() * batch.nstruc) * 1e-10 * (F0 ** 2).sum()
    tol = np.array([2 * np.sqrt(parts[0] * DE) + DE, 2 * np.sqrt(parts[1] * DF) + DF])
    print(f"  loss {loss} vs {parts}; |diff| {np.abs(loss - parts)} tol {tol}", flush=True)
    assert np.all(np.abs(loss - parts) <= tol), (loss, parts, tol)
    err = np.linalg.norm(g - gr) / np.linalg.norm(gr)
    print(f"  grad rel err {err:.2e}", flush=True)
    assert err <= 1e-5, err
