"""The trainer around ann_train_step against the FP64 oracle (through the C
ABI): the reference's energy-loss variants and normalisations
(src/nnp/nnptefr2.cpp:1336-1372, :602-612), per-structure weights, Taylor
augmented copies trained on energies only (:2466-2473), and rotating batches
drawn by the reference's sampler from a training set resident in HBM
(:1309-1315).

Bars: loss terms relative 1e-6; gradient relative (L2) 1e-5.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _check(eng, orc, batch, Eref, Fref, wEs=None, wFs=None, **loss):
    import torch
    dev = "cuda:0"
    lossg, g = eng.train_step(torch.tensor(Eref, device=dev), torch.tensor(Fref, device=dev), 1.0, 0.1, 0.0)
    _, go, _, _, parts = orc.train_loss_grad(batch, Eref, Fref, 1.0, 0.1, wEs=wEs, wFs=wFs, **loss)
    lg = lossg.cpu().numpy()
    assert np.allclose(lg, parts, rtol=1e-6, atol=1e-14), (lg, parts)
    err = float(np.linalg.norm(g.cpu().numpy() - go) / np.linalg.norm(go))
    print(f"  loss {lg} grad rel err {err:.2e}", flush=True)
    assert err <= 1e-5, err


def _labels(orc, model, batch, rng, scale=0.05):
    E0, F0, _ = orc.compute_batch(batch, nthreads=4)
    evac = np.array([sum(model.evac[s] for s in batch.struc(k)[2]) for k in range(batch.nstruc)])
    return E0 - evac + rng.normal(size=batch.nstruc) * scale * batch.natoms, F0 + rng.normal(size=F0.shape) * scale


@pytest.mark.parametrize("loss,norm", [("MAE", "LINEAR"), ("HUBER", "LINEAR"), ("ASINH", "LINEAR"),
                                       ("MSE", "IDENTITY"), ("MSE", "SQRT"), ("MSE", "CBRT"), ("HUBER", "LOG")])
def test_loss_variants(oracle_mod, loss, norm):
    import torch
    from ann_amd import configs
    from ann_amd.gpu import Engine
    model, batch = configs.model_lips(), configs.batch_lips(3, seed=91)
    print(f"{loss}/{norm}", flush=True)
    eng = Engine(model, "cuda:0")
    eng.set_batch(batch)
    eng.neighbors()
    eng.set_loss(loss, 0.3, norm)
    orc = oracle_mod.Oracle(model)
    Eref, Fref = _labels(orc, model, batch, np.random.default_rng(1))
    _check(eng, orc, batch, Eref, Fref, loss=Engine.LOSS[loss], delta=0.3, norm=Engine.NORM[norm])


def test_structure_weights(oracle_mod):
    import torch
    from ann_amd import configs
    from ann_amd.gpu import Engine
    model, batch = configs.model_lips(), configs.batch_lips(3, seed=92)
    eng = Engine(model, "cuda:0")
    eng.set_batch(batch)
    eng.neighbors()
    wE, wF = np.array([1.0, 0.5, 2.0]), np.array([1.0, 0.0, 1.5])
    eng.set_struc_weights(torch.tensor(wE, device="cuda:0"), torch.tensor(wF, device="cuda:0"))
    orc = oracle_mod.Oracle(model)
    Eref, Fref = _labels(orc, model, batch, np.random.default_rng(2))
    _check(eng, orc, batch, Eref, Fref, wEs=wE, wFs=wF)


def test_taylor_copies_train_energy_only(oracle_mod):
    """Augmented copies (no reference forces) enter with force weight 0."""
    import torch
    from ann_amd import configs
    from ann_amd.augment import taylor_augment
    from ann_amd.gpu import Engine
    model = configs.model_water()
    base = configs.batch_water(2, seed=93, nmol=16, L=8.0)
    orc = oracle_mod.Oracle(model)
    E0, F0 = _labels(orc, model, base, np.random.default_rng(3))
    batch, E, F, hasf = taylor_augment(base, E0, F0, nadd=2, rdelta=0.02, seed=4)
    assert batch.nstruc == 6 and hasf.sum() == 2
    eng = Engine(model, "cuda:0")
    eng.set_batch(batch)
    eng.neighbors()
    wF = hasf.astype(np.float64)
    eng.set_struc_weights(None, torch.tensor(wF, device="cuda:0"))
    _check(eng, orc, batch, E, F, wFs=wF)


def test_rotating_batches_from_a_resident_set(oracle_mod):
    """Three steps, each a new batch drawn by the sampler and gathered on the
    GPU; the cached G and moments of one batch must not leak into the next."""
    from ann_amd import configs
    from ann_amd.batch import BatchSampler, ResidentSet
    from ann_amd.gpu import Engine
    model = configs.model_lips()
    full = configs.batch_lips(8, seed=94)
    orc = oracle_mod.Oracle(model)
    Eref, Fref = _labels(orc, model, full, np.random.default_rng(5))
    rs = ResidentSet(full, Eref, Fref)
    sampler = BatchSampler(3, 8, seed=6)
    eng = Engine(model, "cuda:0")
    for step in range(3):
        idx = sampler.next()
        print(f"step {step}: structures {idx.tolist()}", flush=True)
        b = rs.gather(idx)
        eng.set_batch_tensors(b["atom_off"], b["pos"], b["cells"], b["species"])
        eng.neighbors()
        lossg, g = eng.train_step(b["energy"], b["forces"], 1.0, 0.1, 0.0)
        sub = full.subset(idx)
        o = full.atom_off
        Fsub = np.concatenate([Fref[o[k]:o[k + 1]] for k in idx])
        _, go, _, _, parts = orc.train_loss_grad(sub, Eref[idx], Fsub, 1.0, 0.1)
        assert np.allclose(lossg.cpu().numpy(), parts, rtol=1e-6)
        err = float(np.linalg.norm(g.cpu().numpy() - go) / np.linalg.norm(go))
        assert err <= 1e-5, (step, err)


def test_host_set_up_matches_device_set_up(oracle_mod):
    """ann_set_batch_host (offsets, cells and species from the host, no device
    readback in the set-up) against ann_set_batch on the same device arrays:
    bitwise equal energies, forces, losses and gradients over four rotating
    batches with Adam updates between them (batches of different sizes and
    compositions, so the grid, tiles and scratch are rebuilt), and the oracle's
    loss and gradient on the first step."""
    import torch
    from ann_amd import configs
    from ann_amd.batch import BatchSampler, ResidentSet
    from ann_amd.gpu import Engine, FAST
    model = configs.model_lips()
    full = configs.Batch.concat([configs.batch_lips(4, seed=93), configs.batch_lips(4, seed=95, a=5.9)])
    orc = oracle_mod.Oracle(model)
    Eref, Fref = _labels(orc, model, full, np.random.default_rng(7))
    rs = ResidentSet(full, Eref, Fref)
    e1, e2 = Engine(model, "cuda:0"), Engine(model, "cuda:0")
    sizes = [3, 5, 2, 4]
    rng = np.random.default_rng(8)
    for step, k in enumerate(sizes):
        idx = np.sort(rng.choice(full.nstruc, size=k, replace=False))
        print(f"step {step}: structures {idx.tolist()}", flush=True)
        b = rs.gather(idx)
        e1.set_batch_tensors(b["atom_off"], b["pos"], b["cells"], b["species"])
        e2.set_batch_host(b["atom_off"], b["pos"], b["cells_host"], b["species_host"])
        outs = []
        for e in (e1, e2):
            e.neighbors()
            E, F, V = e.energy_forces(FAST)
            lossg, g = e.train_step(b["energy"], b["forces"], 1.0, 0.1, 1e-3)
            outs.append((E.clone(), F.clone(), V.clone(), lossg.clone(), g.clone()))
        torch.cuda.synchronize()
        for a, c in zip(outs[0], outs[1]):
            assert torch.equal(a, c), step
        if step == 0:
            sub = full.subset(idx)
            o = full.atom_off
            Fsub = np.concatenate([Fref[o[q]:o[q + 1]] for q in idx])
            _, go, _, _, parts = orc.train_loss_grad(sub, Eref[idx], Fsub, 1.0, 0.1)
            assert np.allclose(outs[1][3].cpu().numpy(), parts, rtol=1e-6)
            err = float(np.linalg.norm(outs[1][4].cpu().numpy() - go) / np.linalg.norm(go))
            assert err <= 1e-5, err
    assert np.array_equal(e1.get_params(), e2.get_params())


def test_host_set_up_rejects_bad_species(oracle_mod):
    import torch
    from ann_amd import configs
    from ann_amd.gpu import AnnError, Engine
    model = configs.model_lips()
    b = configs.batch_lips(2, seed=96)
    eng = Engine(model, "cuda:0")
    pos = torch.tensor(b.pos, device="cuda:0")
    sp = b.species.copy()
    sp[3] = 7
    with pytest.raises(AnnError):
        eng.set_batch_host(b.atom_off, pos, b.cells, sp)
    eng.set_batch_host(b.atom_off, pos, b.cells, b.species)     # still usable
    eng.neighbors()
    E, F, _ = eng.energy_forces()
    Eo, Fo, _ = oracle_mod.Oracle(model).compute_batch(b, nthreads=4)
    assert np.max(np.abs(E.cpu().numpy() - Eo) / b.natoms) <= 1e-6
