"""GPU parity of every basis family, cutoff and neuron against the FP64 oracle
(through the C ABI), beyond the GAUSSIAN / GAUSS / TANH / SOFTPLUS defaults
of the benchmark workloads.

Bars: G <= 2e-5 x max |G| per structure; energies <= 1e-6 eV/atom; forces
<= 1e-5 relative RMS; basis tangent <= 1e-5 relative RMS; atom energies
<= 1e-6 and dE/dG <= 1e-5 relative RMS; second-backward parameter gradient
<= 1e-5 relative (L2).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

FAMILIES = [("SECH", "GAUSS2", "COS"), ("LOGISTIC", "SECH", "POLY3"), ("TANH", "STUDENT3", "TANH"),
            ("LOGCOSH", "STUDENT4", "COS"), ("LOGCOSH2", "STUDENT5", "STEP"), ("GAUSSIAN", "GAUSS", "POLY3")]
NEURONS = ["LINEAR", "SIGMOID", "TANH", "ISRU", "ARCTAN", "RELU", "ELU", "TANHRE", "SQRE", "SWISH", "GELU",
           "MISH", "PFLU", "SERF", "LOGISH", "SOFTPLUS", "SQPLUS", "ATISH"]


def _batch():
    from ann_amd import configs
    return configs.batch_water(2, seed=71, nmol=16, L=8.0)


def _engine(model, batch):
    from ann_amd.gpu import Engine
    eng = Engine(model, "cuda:0")
    eng.set_batch(batch)
    eng.neighbors()
    return eng


@pytest.mark.parametrize("form", [0, 1])
@pytest.mark.parametrize("rad,ang,cut", FAMILIES)
def test_family_symm_energy_force_tangent(oracle_mod, rad, ang, cut, form):
    import torch
    from ann_amd import configs
    model, batch = configs.model_family(rad, ang, cutoff=cut), _batch()
    print(f"{rad}/{ang}/{cut} form {form}", flush=True)
    eng = _engine(model, batch)
    orc = oracle_mod.Oracle(model)
    G = eng.symm(form).cpu().numpy().astype(np.float64)
    off = batch.atom_off
    rng = np.random.default_rng(3)
    u = rng.normal(size=(batch.ntotal, 3))
    v = eng.tangent(torch.tensor(u, device="cuda:0"), form).cpu().numpy()
    for k in range(batch.nstruc):
        pos, cell, sp = batch.struc(k)
        Gr = orc.symm(pos, cell, sp)
        assert np.abs(G[off[k]:off[k + 1], :Gr.shape[1]] - Gr).max() <= 2e-5 * np.abs(Gr).max(), k
        vr = orc.tangent(pos, cell, sp, u[off[k]:off[k + 1]])
        vg = v[off[k]:off[k + 1], :vr.shape[1]]
        assert np.sqrt(((vg - vr) ** 2).sum() / (vr ** 2).sum()) <= 1e-5, k
    E, F, _ = eng.energy_forces(form)
    Eo, Fo, _ = orc.compute_batch(batch, nthreads=4)
    de = np.max(np.abs(E.cpu().numpy() - Eo) / batch.natoms)
    rf = np.sqrt(((F.cpu().numpy() - Fo) ** 2).sum() / (Fo ** 2).sum())
    print(f"  dE/atom={de:.2e} force relRMS={rf:.2e}", flush=True)
    assert de <= 1e-6 and rf <= 1e-5


@pytest.mark.parametrize("neuron", NEURONS)
def test_neuron_forward_and_second_backward(oracle_mod, neuron):
    import torch
    from ann_amd import configs
    model, batch = configs.model_family(neuron=neuron), _batch()
    print(f"neuron {neuron}", flush=True)
    eng = _engine(model, batch)
    orc = oracle_mod.Oracle(model)
    for form in (0, 1):
        G = eng.symm(form)
        E, dE = eng.net(G, form)
        Gc = G.cpu().numpy().astype(np.float64)[:, :orc.ld]
        Eo, dEo = orc.net(batch.species, Gc)
        assert np.max(np.abs(E.cpu().numpy() - Eo)) <= 1e-6 * max(1.0, np.abs(Eo).max()), form
        d = dE.cpu().numpy()[:, :orc.ld]
        assert np.sqrt(((d - dEo) ** 2).sum() / (dEo ** 2).sum()) <= 1e-5, form
    # second backward: sum_i d/dp [gE E_i + dE_i/dG . v_i]
    G = eng.symm(1)
    rng = np.random.default_rng(8)
    v = (rng.normal(size=(batch.ntotal, eng.ldg)) * 0.1).astype(np.float32)
    gE = rng.normal(size=batch.nstruc)
    g = eng.net_bwd2(G, torch.tensor(v, device="cuda:0"), torch.tensor(gE, device="cuda:0")).cpu().numpy()
    Gc = G.cpu().numpy().astype(np.float64)
    poff = np.concatenate([[0], np.cumsum([orc.nparams(a) for a in range(model.nspecies)])])
    ref = np.zeros(poff[-1])
    sof = np.repeat(np.arange(batch.nstruc), batch.natoms)
    for i in range(batch.ntotal):
        a = batch.species[i]
        ref[poff[a]:poff[a + 1]] += orc.nn_grad_rop(a, Gc[i, :orc.ld], v[i, :orc.ld].astype(np.float64),
                                                    gE[sof[i]])
    err = np.linalg.norm(g - ref) / np.linalg.norm(ref)
    print(f"  second backward rel err {err:.2e}", flush=True)
    assert err <= 1e-5
