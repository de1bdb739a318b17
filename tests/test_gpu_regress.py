"""Regression tests of fixed defects, run through the C ABI against the FP64
oracle (bars as in test_gpu_parity: G <= 2e-5 x max |G|, P relative RMS
<= 1e-5, forces relative RMS <= 1e-5, energies <= 1e-6 eV/atom).

* The moment form's per-atom moment store is sized by the fused calls; a
  forward call on a larger batch before any fused call must not use it.
* A captured energy/force graph must not replay a species permutation of an
  earlier batch when ann_set_batch keeps the same buffers but changes the
  species composition.
* A neighbour build error is reported with its own status code and does not
  stick to later builds.
"""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_forward_on_grown_batch_after_fused_call(oracle_mod):
    import torch
    from ann_amd import configs
    from ann_amd.gpu import Engine, FAST
    model = configs.model_water()
    small, big = configs.batch_water(1, seed=61), configs.batch_water(3, seed=62)
    eng = Engine(model, "cuda:0")
    eng.set_batch(small)
    eng.neighbors()
    eng.energy_forces(FAST)            # sizes the work buffers (and the moment store) for 192 atoms
    torch.cuda.synchronize()
    eng.set_batch(big)                 # 576 atoms
    eng.neighbors()
    print("grown batch: symm, net, symm_bwd", flush=True)
    G = eng.symm(FAST)
    _, dE = eng.net(G, FAST)
    P = eng.symm_bwd(dE, FAST).cpu().numpy()
    Gc = G.cpu().numpy().astype(np.float64)
    dEc = dE.cpu().numpy().astype(np.float64)
    orc = oracle_mod.Oracle(model)
    off = big.atom_off
    nl = eng.nlist()
    for k in range(big.nstruc):
        pos, cell, sp = big.struc(k)
        Gr = orc.symm(pos, cell, sp)
        assert np.abs(Gc[off[k]:off[k + 1], :Gr.shape[1]] - Gr).max() <= 2e-5 * np.abs(Gr).max()
        _, _, Pr = orc.force(pos, cell, sp, dEc[off[k]:off[k + 1], :orc.ld])
        s0, s1 = nl["off"][off[k]], nl["off"][off[k + 1]]
        assert np.sqrt(((P[s0:s1] - Pr) ** 2).sum() / (Pr ** 2).sum()) <= 1e-5
    # and the fused call on the grown batch (it grows the store)
    E, F, _ = eng.energy_forces(FAST)
    Eo, Fo, _ = orc.compute_batch(big, nthreads=4)
    assert np.sqrt(((F.cpu().numpy() - Fo) ** 2).sum() / (Fo ** 2).sum()) <= 1e-5


def test_ef_graph_follows_species_composition(oracle_mod):
    import torch
    from ann_amd import configs
    from ann_amd.gpu import Engine, FAST, _check, lib
    model = configs.model_water()
    batch = configs.batch_water(2, seed=63)
    eng = Engine(model, "cuda:0")
    eng.ef_graph(True)
    eng.set_batch(batch)
    eng.neighbors()
    out = (torch.empty(batch.nstruc, dtype=torch.float64, device="cuda:0"),
           torch.empty((batch.ntotal, 3), dtype=torch.float64, device="cuda:0"),
           torch.empty((batch.nstruc, 3, 3), dtype=torch.float64, device="cuda:0"))
    eng.energy_forces(FAST, out=out)   # eager
    eng.energy_forces(FAST, out=out)   # captured and replayed
    # a different composition in the same buffers: every third H becomes O
    sp = batch.species.copy()
    h = np.flatnonzero(sp == 0)
    sp[h[::3]] = 1
    eng.d_species.copy_(torch.from_numpy(sp.astype(np.int32)).to("cuda:0"))
    torch.cuda.synchronize()
    _check(lib().ann_set_batch(eng.h, C.byref(eng._bdesc)))
    eng.neighbors()
    print("composition changed in place; replay", flush=True)
    E, F, _ = eng.energy_forces(FAST, out=out)
    b2 = configs.Batch(batch.natoms, batch.pos, batch.cells, sp)
    Eo, Fo, _ = oracle_mod.Oracle(model).compute_batch(b2, nthreads=4)
    assert np.max(np.abs(E.cpu().numpy() - Eo) / b2.natoms) <= 1e-6
    assert np.sqrt(((F.cpu().numpy() - Fo) ** 2).sum() / (Fo ** 2).sum()) <= 1e-5


def test_neighbour_errors_have_distinct_codes_and_do_not_stick():
    from ann_amd import configs
    from ann_amd.gpu import AnnError, Engine
    model = configs.model_water()
    rng = np.random.default_rng(3)
    crowded = configs.single(rng.uniform(0, 5.0, size=(600, 3)), np.eye(3) * 40.0, np.zeros(600, np.int32))
    eng = Engine(model, "cuda:0")
    eng.set_batch(crowded)
    with pytest.raises(AnnError) as ei:
        eng.neighbors()
    assert ei.value.code == -3 and "capacity" in str(ei.value)
    # a cell far narrower than the cutoff exceeds the image limit: unsupported, not capacity
    thin = configs.single([[0.1, 0.1, 0.1]], np.diag([0.25, 30.0, 30.0]), [0])
    eng.set_batch(thin)
    with pytest.raises(AnnError) as ei:
        eng.neighbors()
    assert ei.value.code == -2
    # the same context builds a valid list afterwards
    good = configs.batch_water(1, seed=64)
    eng.set_batch(good)
    eng.neighbors()
    assert eng.nlist_size()[0] > 0
