"""Training restart (ann_get/set_optimizer_state + ann_amd/restart.py, the
state of the trainer's restart file, src/nnp/nnptefr2.cpp:73-195, 501-566):
three Adam steps, a checkpoint, three more steps; a fresh engine loading the
checkpoint and taking the same three steps ends with bitwise the same
parameters and Adam state (training steps are deterministic), and the input
scaling is restored.  Negative control: the same resume without the Adam
moments ends elsewhere.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _setup():
    from ann_amd import configs
    from ann_amd.gpu import Engine
    model, batch = configs.model_lips(), configs.batch_lips(3, seed=44)
    eng = Engine(model, "cuda:0")
    eng.set_batch(batch)
    eng.neighbors()
    return batch, eng


def _steps(eng, Eref, Fref, n):
    from ann_amd.gpu import FAST
    for _ in range(n):
        eng.train_step(Eref, Fref, 1.0, 0.1, 1e-3, FAST)


def test_resume_matches_uninterrupted(tmp_path):
    import torch
    from ann_amd.precond import apply_preconditioning
    from ann_amd.restart import load_restart, save_restart
    print("case lips3 seed 44 restart", flush=True)
    batch, a = _setup()
    apply_preconditioning(a, "MEAN", "DEV")
    rng = np.random.default_rng(13)
    Eref = torch.tensor(rng.normal(size=batch.nstruc), device="cuda:0")
    Fref = torch.tensor(rng.normal(size=(batch.ntotal, 3)), device="cuda:0")
    _steps(a, Eref, Fref, 3)
    path = str(tmp_path / "ckpt.npz")
    save_restart(path, a, iteration=3)
    _steps(a, Eref, Fref, 3)
    pa, (ma, va, ta) = a.get_params(), a.get_optimizer_state()
    assert ta == 6

    _, b = _setup()
    assert load_restart(path, b) == 3
    for s in range(b.nspecies):
        wa, ba, _, _ = a.scaling(s)
        wb, bb, _, _ = b.scaling(s)
        assert np.array_equal(wa, wb) and np.array_equal(ba, bb)
    _steps(b, Eref, Fref, 3)
    pb, (mb, vb, tb) = b.get_params(), b.get_optimizer_state()
    assert tb == 6
    assert np.array_equal(pa, pb) and np.array_equal(ma, mb) and np.array_equal(va, vb)

    _, c = _setup()
    load_restart(path, c)
    c.set_optimizer_state(np.zeros(c.nparams), np.zeros(c.nparams), 0)
    _steps(c, Eref, Fref, 3)
    assert np.abs(c.get_params() - pa).max() > 1e-6
