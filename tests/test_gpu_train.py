"""GPU parity of the training path against the FP64 oracle.

Tolerances: basis tangent relative RMS <= 1e-5; second-backward parameter
gradient relative (L2) error <= 1e-5; loss terms relative <= 1e-6; Adam
update against the oracle's Adam on the same gradient to 1e-12.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _setup(name):
    from ann_amd import configs
    from ann_amd.gpu import Engine
    if name == "lips3":
        model, batch = configs.model_lips(), configs.batch_lips(3, seed=41)
    else:
        model, batch = configs.model_water(), configs.batch_water(2, seed=42)
    eng = Engine(model, "cuda:0")
    eng.set_batch(batch)
    eng.neighbors()
    return model, batch, eng


@pytest.mark.parametrize("form", [0, 1, 2])
@pytest.mark.parametrize("name", ["lips3", "water2"])
def test_tangent_parity(oracle_mod, name, form):
    import torch
    model, batch, eng = _setup(name)
    print(f"case {name} form {form}", flush=True)
    rng = np.random.default_rng(7)
    u = rng.normal(size=(batch.ntotal, 3))
    v = eng.tangent(torch.tensor(u, device="cuda:0"), form).cpu().numpy()
    orc = oracle_mod.Oracle(model)
    off = batch.atom_off
    for k in range(batch.nstruc):
        pos, cell, sp = batch.struc(k)
        vr = orc.tangent(pos, cell, sp, u[off[k]:off[k + 1]])
        vg = v[off[k]:off[k + 1], :vr.shape[1]]
        err = np.sqrt(((vg - vr) ** 2).sum() / (vr ** 2).sum())
        assert err <= 1e-5, (k, err)


@pytest.mark.parametrize("name", ["lips3", "water2"])
def test_second_backward_parity(oracle_mod, name):
    import torch
    model, batch, eng = _setup(name)
    print(f"case {name}", flush=True)
    G = eng.symm(1)
    rng = np.random.default_rng(8)
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
@pytest.mark.parametrize("name", ["lips3", "water2"])
def test_train_step_parity(oracle_mod, name, form):
    import torch
    model, batch, eng = _setup(name)
    print(f"case {name} form {form}", flush=True)
    rng = np.random.default_rng(9)
    orc = oracle_mod.Oracle(model)
    E0, F0, _ = orc.compute_batch(batch, nthreads=4)
    evac = np.array([sum(model.evac[s] for s in batch.struc(k)[2]) for k in range(batch.nstruc)])
    Eref = E0 - evac + rng.normal(size=batch.nstruc) * 0.05 * batch.natoms
    Fref = F0 + rng.normal(size=F0.shape) * 0.05
    loss, g = eng.train_step(torch.tensor(Eref, device="cuda:0"), torch.tensor(Fref, device="cuda:0"),
                             1.0, 0.1, 0.0, form)
    loss, g = loss.cpu().numpy(), g.cpu().numpy()
    L, gr, _, _, parts = orc.train_loss_grad(batch, Eref, Fref, 1.0, 0.1)
    assert np.allclose(loss, parts, rtol=1e-6, atol=0), (loss, parts)
    err = np.linalg.norm(g - gr) / np.linalg.norm(gr)
    print(f"  loss {loss} vs {parts}; grad rel err {err:.2e}", flush=True)
    assert err <= 1e-5, err


def test_adam_update_matches_oracle(oracle_mod):
    import torch
    model, batch, eng = _setup("lips3")
    rng = np.random.default_rng(10)
    Eref = torch.tensor(rng.normal(size=batch.nstruc), device="cuda:0")
    Fref = torch.tensor(rng.normal(size=(batch.ntotal, 3)), device="cuda:0")
    p0 = eng.get_params()
    m1 = np.zeros_like(p0)
    m2 = np.zeros_like(p0)
    p = p0.copy()
    for t in range(1, 4):
        _, g = eng.train_step(Eref, Fref, 1.0, 0.1, 1e-3, 1)
        p = oracle_mod.adam_step(p, m1, m2, g.cpu().numpy(), 1e-3, 1e-16, t)
    pg = eng.get_params()
    assert np.allclose(pg, p, rtol=0, atol=1e-12)
    assert not np.allclose(pg, p0)


def test_train_graph_replays_match_eager():
    """Six Adam steps eagerly and through the captured graph (replays with
    eager work and synchronisations between them): bitwise identical losses,
    gradients and parameters after every step."""
    import torch
    from ann_amd.gpu import FAST
    model, batch, e1 = _setup("lips3")
    _, _, e2 = _setup("lips3")
    e2.train_graph(True)
    rng = np.random.default_rng(12)
    Eref = torch.tensor(rng.normal(size=batch.nstruc), device="cuda:0")
    Fref = torch.tensor(rng.normal(size=(batch.ntotal, 3)), device="cuda:0")
    o1 = (torch.empty(2, dtype=torch.float64, device="cuda:0"),
          torch.empty(e1.nparams, dtype=torch.float64, device="cuda:0"))
    o2 = (torch.empty(2, dtype=torch.float64, device="cuda:0"),
          torch.empty(e2.nparams, dtype=torch.float64, device="cuda:0"))
    for step in range(6):
        e1.train_step(Eref, Fref, 1.0, 0.1, 1e-3, FAST, out=o1)
        e2.train_step(Eref, Fref, 1.0, 0.1, 1e-3, FAST, out=o2)
        torch.cuda.synchronize()
        assert torch.equal(o1[0], o2[0]), step
        assert torch.equal(o1[1], o2[1]), step
        assert np.array_equal(e1.get_params(), e2.get_params()), step
