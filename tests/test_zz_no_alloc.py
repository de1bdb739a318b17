"""The allocation-free force reduction (ann_force_reduce_noalloc, static
per-GPU scratch) against ann_force_reduce on the same per-slot adjoint and
atom energies, through the C ABI.  Both run the same kernels in the same
order, so forces, energies and virials must agree bit for bit; the calls
come in a row on one context, each later one with a smaller batch than the
one before (a 6,912-atom structure that spans two 4,096-atom sum chunks,
then 3 and 1 water structures)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_noalloc_force_reduce_matches_the_allocating_one():
    import torch
    from ann_amd import configs
    from ann_amd.gpu import AnnError, Engine, FAST, _check, _ptr, lib
    model = configs.model_water()
    eng = Engine(model, "cuda:0")
    batches = [configs.struc_water_md(reps=3), configs.batch_water(3, seed=81), configs.batch_water(1, seed=82)]
    for b in batches:
        print(f"noalloc: {b.nstruc} structures, {b.ntotal} atoms", flush=True)
        eng.set_batch(b)
        eng.neighbors()
        G = eng.symm(FAST)
        Ea, dE = eng.net(G, FAST)
        P = eng.symm_bwd(dE, FAST)
        E1, F1, V1 = eng.force_reduce(P, Ea)
        E2, F2, V2 = eng.force_reduce_noalloc(P, Ea)
        torch.cuda.synchronize()
        assert torch.equal(E1, E2) and torch.equal(F1, F2) and torch.equal(V1, V2)
        # without the virial (no per-atom rows in the scratch; the force gather
        # then reduces 4 values per atom instead of 16, in its own fixed order)
        F4, E4 = torch.full_like(F1, np.nan), torch.full_like(E1, np.nan)
        _check(lib().ann_force_reduce(eng.h, _ptr(P), _ptr(Ea), 1, _ptr(F4), _ptr(E4), _ptr(None)))
        F3, E3 = torch.full_like(F1, np.nan), torch.full_like(E1, np.nan)
        eng.force_reduce_noalloc(P, Ea, out=(E3, F3, V2), virial=False)
        torch.cuda.synchronize()
        assert torch.equal(E4, E3) and torch.equal(F4, F3) and torch.equal(E1, E3)
    with pytest.raises(AnnError) as ei:
        _check(lib().ann_force_reduce_noalloc(eng.h, _ptr(None), _ptr(Ea), 1, _ptr(F1), _ptr(E1), _ptr(V1)))
    assert ei.value.code == -1
