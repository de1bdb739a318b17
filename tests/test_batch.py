"""The training-batch sampler (Batch, src/ml/batch.hpp:18-55, as used by
src/nnp/nnptefr2.cpp:676-678 and 1309-1315)."""
import numpy as np
import pytest

from ann_amd.batch import BatchSampler


def test_epochs_partition_the_set_when_size_divides_capacity():
    s = BatchSampler(4, 12, seed=3)
    for _ in range(5):
        seen = []
        for _ in range(3):
            b = s.next()
            assert np.all(np.diff(b) > 0)          # sorted, no repeats
            seen.extend(b.tolist())
        assert sorted(seen) == list(range(12))
        assert s.count == 0


def test_wraparound_batch_then_reshuffle():
    """With 5 of 7, the second draw wraps: indices 5, 6 of the current
    permutation, then 0, 1, 2 of the same permutation; then it reshuffles."""
    s = BatchSampler(5, 7, seed=1)
    perm = s.data.copy()
    b1 = s.next()
    assert sorted(b1.tolist()) == sorted(perm[:5].tolist()) and s.count == 5
    b2 = s.next()
    assert sorted(b2.tolist()) == sorted(perm[[5, 6, 0, 1, 2]].tolist())
    assert s.count == 0 and sorted(s.data.tolist()) == list(range(7))


def test_reference_argument_checks():
    with pytest.raises(ValueError, match="smaller than capacity"):
        BatchSampler(8, 4)
    with pytest.raises(ValueError, match="negative"):
        BatchSampler(-1, 4)
    s = BatchSampler(0, 3)
    assert len(s.next()) == 0


class _OracleEngine:
    """Engine stand-in for the host side of validation_error: the energies of
    the batch it is given come from the FP64 oracle (no vacuum energies)."""

    def __init__(self, orc, model):
        self.orc, self.model = orc, model

    def set_batch_host(self, atom_off, pos, cells, species):
        from ann_amd.configs import Batch
        self.b = Batch(np.diff(atom_off), pos.numpy(), np.asarray(cells).reshape(-1, 3, 3), np.asarray(species))

    def neighbors(self):
        pass

    def energies(self, form, with_evac=True):
        import torch
        E, _, _ = self.orc.compute_batch(self.b, nthreads=2)
        ev = np.array([sum(self.model.evac[s] for s in self.b.struc(k)[2]) for k in range(self.b.nstruc)])
        return torch.tensor(E - (0.0 if with_evac else ev))


@pytest.mark.parametrize("loss,norm", [("MSE", "LINEAR"), ("MAE", "IDENTITY"), ("HUBER", "SQRT"),
                                       ("ASINH", "CBRT"), ("HUBER", "LOG")])
def test_validation_error_formulas_against_oracle(oracle_mod, loss, norm):
    """The report's loss (src/nnp/nnptefr2.cpp:1375-1411, averaged at :986-991)
    on the oracle's own energies equals the oracle's energy loss with zero force
    weight, and its RMSE the mean squared per-atom error, to FP64 rounding."""
    from ann_amd import configs
    from ann_amd.batch import ResidentSet, validation_error
    from ann_amd.gpu import Engine
    model = configs.model_lips()
    full = configs.batch_lips(4, seed=98)
    orc = oracle_mod.Oracle(model)
    E0, F0, _ = orc.compute_batch(full, nthreads=2)
    evac = np.array([sum(model.evac[s] for s in full.struc(k)[2]) for k in range(full.nstruc)])
    Eref = E0 - evac + np.random.default_rng(99).normal(size=full.nstruc) * 0.05 * full.natoms
    rs = ResidentSet(full, Eref, np.zeros_like(F0), device="cpu")
    idx = np.array([3, 0, 2])
    v = validation_error(_OracleEngine(orc, model), rs, idx, loss, 0.3, norm, chunk=2)
    sub = full.subset(idx)
    _, _, Eo, _, parts = orc.train_loss_grad(sub, Eref[idx], np.zeros((sub.ntotal, 3)), 1.0, 0.0,
                                             loss=Engine.LOSS[loss], delta=0.3, norm=Engine.NORM[norm])
    assert np.isclose(v["loss"], parts[0], rtol=1e-12, atol=0), (v, parts)
    assert np.isclose(v["rmse"], np.sqrt(np.mean(((Eo - Eref[idx]) / sub.natoms) ** 2)), rtol=1e-12, atol=0)
