"""Checkpoint file of ann_amd/restart.py on the host (CPU only, a stand-in
engine): every array survives the round trip bit for bit, the order of
restoration puts the Adam state after the parameters (setting parameters
clears it), and files for another network or version are refused.
"""
import numpy as np
import pytest

from ann_amd.restart import load_restart, save_restart


class _Engine:
    def __init__(self, nspecies=2, nparams=7, seed=0):
        rng = np.random.default_rng(seed)
        self.nspecies, self.nparams = nspecies, nparams
        self.p = rng.normal(size=nparams)
        self.m, self.v, self.t = rng.normal(size=nparams), rng.random(nparams), int(rng.integers(1, 100))
        self.sc = {a: (rng.normal(size=3 + a), rng.normal(size=3 + a), float(rng.normal()), float(rng.normal()))
                   for a in range(nspecies)}

    def get_params(self):
        return self.p.copy()

    def set_params(self, p):
        self.p = np.array(p)
        self.m, self.v, self.t = np.zeros(self.nparams), np.zeros(self.nparams), 0

    def get_optimizer_state(self):
        return self.m.copy(), self.v.copy(), self.t

    def set_optimizer_state(self, m, v, t):
        self.m, self.v, self.t = np.array(m), np.array(v), int(t)

    def scaling(self, a):
        return self.sc[a]

    def set_scaling(self, a, w, b, ow, ob):
        self.sc[a] = (np.array(w), np.array(b), ow, ob)


def test_round_trip(tmp_path):
    src, dst = _Engine(seed=1), _Engine(seed=2)
    path = str(tmp_path / "r.npz")
    save_restart(path, src, iteration=17)
    assert load_restart(path, dst) == 17
    assert np.array_equal(dst.p, src.p) and np.array_equal(dst.m, src.m) and np.array_equal(dst.v, src.v)
    assert dst.t == src.t
    for a in range(src.nspecies):
        for x, y in zip(src.sc[a], dst.sc[a]):
            assert np.array_equal(x, y)


def test_other_network_refused(tmp_path):
    path = str(tmp_path / "r.npz")
    save_restart(path, _Engine(nparams=7))
    with pytest.raises(ValueError):
        load_restart(path, _Engine(nparams=8))
    with pytest.raises(ValueError):
        load_restart(path, _Engine(nspecies=3))
