"""Input preconditioning statistics on the GPU (ann_input_stats, restating the
loops of src/nnp/nnptefr2.cpp:690-787) and the preconditioning they feed
(ann_amd/precond.py, :788-867).

Tolerances:
  against numpy statistics of the same FP32 G rows in FP64: min and max exact,
  harmonic mean, mean, deviation (N - 1) and rms within 1e-12 relative;
  against the FP64 oracle's G: min, max, mean, deviation and rms within
  5e-5 x max |G| of the input (the G parity bound of the FP32 rows);
  after MEAN / DEV preconditioning: every scaled input has mean 0 (1e-9) and
  deviation 1 (1e-9) over its species, and energies match the oracle given the
  same input scaling within 1e-6 eV/atom.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

DBL_MAX = np.finfo(np.float64).max


def _engine(model, batch):
    from ann_amd.gpu import Engine
    eng = Engine(model, "cuda:0")
    eng.set_batch(batch)
    eng.neighbors()
    return eng


def _numpy_stats(G, species, nin, S):
    out = []
    for a in range(S):
        rows = G[species == a, :nin[a]]
        n = len(rows)
        with np.errstate(divide="ignore", invalid="ignore"):
            out.append((n, rows.min(0), rows.max(0), n / (1.0 / rows).sum(0), rows.mean(0),
                        rows.std(0, ddof=1), np.sqrt((rows * rows).mean(0))))
    return out


@pytest.mark.parametrize("name", ["lips3", "water2"])
def test_stats_match_numpy_and_oracle(oracle_mod, name):
    from ann_amd import configs
    if name == "lips3":
        model, batch = configs.model_lips(), configs.batch_lips(3, seed=41)
    else:
        model, batch = configs.model_water(), configs.batch_water(2, seed=45)
    print(f"case {name}", flush=True)
    eng = _engine(model, batch)
    G = eng.symm().cpu().numpy().astype(np.float64)
    cnt, st = eng.input_stats()
    nin = [model.ninput(a) for a in range(eng.nspecies)]
    ref = _numpy_stats(G, batch.species, nin, eng.nspecies)
    orc = oracle_mod.Oracle(model)
    Go = np.concatenate([orc.symm(*batch.struc(k)) for k in range(batch.nstruc)])
    oref = _numpy_stats(Go, batch.species, nin, eng.nspecies)
    for a in range(eng.nspecies):
        n = nin[a]
        assert cnt[a] == ref[a][0] == np.sum(batch.species == a)
        assert np.array_equal(st[a, 0, :n], ref[a][1]) and np.array_equal(st[a, 1, :n], ref[a][2])
        for q in (2, 3, 4, 5):
            assert np.allclose(st[a, q, :n], ref[a][q + 1], rtol=1e-12, atol=0), (a, q)
        assert not np.any(st[a, :, n:])
        scale = np.abs(Go[batch.species == a, :n]).max(0) + 1e-30
        for q in (0, 1, 3, 4, 5):
            assert np.all(np.abs(st[a, q, :n] - oref[a][q + 1]) <= 5e-5 * scale), (a, q)


def test_species_without_atoms():
    from ann_amd import configs
    model = configs.model_lips()
    b = configs.batch_lips(2, seed=42)
    sp = np.where(b.species == 1, 2, b.species).astype(np.int32)
    batch = configs.Batch(b.natoms, b.pos, b.cells, sp)
    print("case lips2 without species 1", flush=True)
    eng = _engine(model, batch)
    cnt, st = eng.input_stats()
    n = model.ninput(1)
    assert cnt[1] == 0 and cnt[0] + cnt[2] == batch.ntotal
    assert np.all(st[1, 0, :n] == DBL_MAX) and np.all(st[1, 1, :n] == -DBL_MAX)
    assert np.all(np.isnan(st[1, 3, :n])) and np.all(np.isnan(st[1, 5, :n]))
    assert np.all(np.isfinite(st[0, 3, :model.ninput(0)]))


def test_preconditioning_standardises_inputs(oracle_mod):
    from ann_amd import configs
    from ann_amd.gpu import FAST
    from ann_amd.precond import apply_preconditioning
    model, batch = configs.model_lips(), configs.batch_lips(3, seed=43)
    print("case lips3 seed 43 MEAN/DEV", flush=True)
    eng = _engine(model, batch)
    G = eng.symm().cpu().numpy().astype(np.float64)
    pre = apply_preconditioning(eng, "MEAN", "DEV")
    for a, (w, b) in enumerate(pre):
        x = (G[batch.species == a, :len(w)] + b) * w
        live = x.std(0, ddof=1) > 0
        assert np.all(np.abs(x.mean(0)) <= 1e-9 * np.maximum(1.0, np.abs(w)))
        assert np.allclose(x.std(0, ddof=1)[live], 1.0, rtol=0, atol=1e-9)
    m2 = model.copy()
    for a, (w, b) in enumerate(pre):
        m2.nets[a].inpw, m2.nets[a].inpb, m2.nets[a].outw, m2.nets[a].outb = w, b, 1.0, 0.0
    Eo, _, _ = oracle_mod.Oracle(m2).compute_batch(batch, nthreads=4)
    E, _, _ = eng.energy_forces(FAST)
    E = E.cpu().numpy()
    assert np.all(np.abs(E - np.asarray(Eo)) / batch.natoms <= 1e-6), np.abs(E - Eo) / batch.natoms
