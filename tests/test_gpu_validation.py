"""The trainer's validation error (ann_amd.batch.validation_error: the
validation loss and RMSE of src/nnp/nnptefr2.cpp:1375-1411, averaged at
:986-991) on structures of a resident set, against the FP64 oracle.

The oracle's training loss with zero force weight is the same energy loss
averaged over the structures (oracle/nnp.cpp, orc_train_loss_grad_ex), so it
is the checker for every loss and normalisation; with MSE and LINEAR
normalisation 2 * loss is the mean squared per-atom energy error, which pins
the RMSE.  Bar: the change the energy parity bar (1e-6 eV/atom) allows, per
structure |d dU| <= 1e-6 n norm(n) = e, and every loss has |dl/dU| <=
max(|dU|, 1) and |d2l/dU2| <= 1, so |dL| <= e mean(max(|dU|, 1)) + e^2 (the
bound is written in the test).  Also: the GPU energies of the report equal
the training step's (same energy loss to 1e-9 relative), chunking does not
change the result, and an empty set is refused.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

NORM = {"IDENTITY": lambda n: np.ones_like(n), "LINEAR": lambda n: 1.0 / n, "SQRT": lambda n: 1.0 / np.sqrt(n),
        "CBRT": lambda n: 1.0 / np.cbrt(n), "LOG": lambda n: 1.0 / (np.log(n) + 1.0)}


def _setup(oracle_mod):
    from ann_amd import configs
    from ann_amd.batch import ResidentSet
    from ann_amd.gpu import Engine
    model = configs.model_lips()
    full = configs.Batch.concat([configs.batch_lips(3, seed=95), configs.batch_lips(3, seed=96, a=5.9)])
    orc = oracle_mod.Oracle(model)
    E0, F0, _ = orc.compute_batch(full, nthreads=4)
    evac = np.array([sum(model.evac[s] for s in full.struc(k)[2]) for k in range(full.nstruc)])
    rng = np.random.default_rng(97)
    Eref = E0 - evac + rng.normal(size=full.nstruc) * 0.05 * full.natoms
    rs = ResidentSet(full, Eref, np.zeros_like(F0))
    eng = Engine(model, "cuda:0")
    return model, full, orc, Eref, rs, eng


@pytest.mark.parametrize("loss,norm", [("MSE", "LINEAR"), ("MAE", "SQRT"), ("HUBER", "IDENTITY"),
                                       ("ASINH", "CBRT"), ("MSE", "LOG")])
def test_validation_error_against_oracle(oracle_mod, loss, norm):
    from ann_amd.batch import validation_error
    from ann_amd.gpu import Engine
    model, full, orc, Eref, rs, eng = _setup(oracle_mod)
    idx = np.array([4, 1, 3, 5])
    print(f"{loss}/{norm}", flush=True)
    v = validation_error(eng, rs, idx, loss, 0.3, norm, chunk=3)
    sub = full.subset(idx)
    _, _, Eo, _, parts = orc.train_loss_grad(sub, Eref[idx], np.zeros((sub.ntotal, 3)), 1.0, 0.0,
                                             loss=Engine.LOSS[loss], delta=0.3, norm=Engine.NORM[norm])
    n = sub.natoms.astype(np.float64)
    e = 1e-6 * np.max(n * NORM[norm](n))
    dU = (Eo - Eref[idx]) * NORM[norm](n)
    tol = e * np.mean(np.maximum(np.abs(dU), 1.0)) + e * e
    print(f"  loss {v['loss']:.12e} oracle {parts[0]:.12e} tol {tol:.2e}", flush=True)
    assert abs(v["loss"] - parts[0]) <= tol
    assert v["structures"] == 4
    # chunking does not change the report
    w = validation_error(eng, rs, idx, loss, 0.3, norm, chunk=512)
    assert np.isclose(w["loss"], v["loss"], rtol=1e-12, atol=0) and np.isclose(w["rmse"], v["rmse"], rtol=1e-12)


def test_validation_rmse_and_training_loss(oracle_mod):
    """RMSE against the oracle (2 x MSE/LINEAR loss = mean squared per-atom
    error); the report's energy loss equals the training step's on the same
    structures (the same GPU energies)."""
    import torch
    from ann_amd.batch import validation_error
    model, full, orc, Eref, rs, eng = _setup(oracle_mod)
    idx = np.arange(full.nstruc)
    v = validation_error(eng, rs, idx, "MSE", 1.0, "LINEAR")
    _, _, Eo, _, parts = orc.train_loss_grad(full, Eref, np.zeros((full.ntotal, 3)), 1.0, 0.0)
    rmse_o = np.sqrt(2.0 * parts[0])
    n = full.natoms.astype(np.float64)
    # per-atom errors off by <= 1e-6 move the RMSE by at most 1e-6
    assert abs(v["rmse"] - rmse_o) <= 1e-6, (v["rmse"], rmse_o)
    assert np.isclose(v["rmse"], np.sqrt(np.mean(((Eo - Eref) / n) ** 2)), rtol=0, atol=1e-6)
    # the training step on the whole set (engine now holds the last chunk = the whole set)
    eng.set_loss("MSE", 1.0, "LINEAR")
    b = rs.gather(idx)
    eng.set_batch_tensors(b["atom_off"], b["pos"], b["cells"], b["species"])
    eng.neighbors()
    lossg, _ = eng.train_step(b["energy"], torch.zeros((full.ntotal, 3), dtype=torch.float64, device="cuda:0"),
                              1.0, 0.0, 0.0)
    assert np.isclose(lossg.cpu().numpy()[0], v["loss"], rtol=1e-9, atol=0), (lossg, v)


def test_empty_validation_set_refused(oracle_mod):
    from ann_amd.batch import validation_error
    model, full, orc, Eref, rs, eng = _setup(oracle_mod)
    with pytest.raises(ValueError):
        validation_error(eng, rs, [])
