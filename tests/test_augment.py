"""Taylor augmentation (src/nnp/nnptefr2.cpp:2383-2482 restated in
ann_amd/augment.py), pinned on the CPU:

* layout: each original followed by its nadd copies, copies without forces,
  structures without forces not copied (a wrong insertion order or count);
* E' = E - sum dr . F exactly, recomputed from the copies' positions (a sign or
  a dropped atom in the Taylor term);
* against the FP64 oracle the estimate is first-order: E(r + dr) - E' shrinks
  fourfold when the displacement halves, and is far below the linear term
  (a displacement applied with the wrong sign or to the wrong atom breaks this);
* directions uniform on the sphere (a biased theta or phi draw).
"""
import numpy as np
import pytest

from ann_amd import configs
from ann_amd.augment import taylor_augment


def _labelled(oracle_mod, nstruc=2, seed=61):
    model = configs.model_lips()
    batch = configs.batch_lips(nstruc, seed=seed)
    orc = oracle_mod.Oracle(model)
    E, F, _ = orc.compute_batch(batch, nthreads=4)
    return model, batch, orc, np.asarray(E), np.asarray(F)


def test_layout_and_linear_term(oracle_mod):
    _, batch, _, E, F = _labelled(oracle_mod)
    nadd = 3
    out, E2, F2, hasf = taylor_augment(batch, E, F, nadd, 0.05, seed=1)
    assert out.nstruc == batch.nstruc * (1 + nadd)
    assert list(hasf) == ([True] + [False] * nadd) * batch.nstruc
    o, oo = batch.atom_off, out.atom_off
    for k in range(batch.nstruc):
        base = k * (1 + nadd)
        p0, c0, s0 = batch.struc(k)
        f0 = F[o[k]:o[k + 1]]
        assert np.array_equal(out.struc(base)[0], p0) and E2[base] == E[k]
        assert np.array_equal(F2[oo[base]:oo[base + 1]], f0)
        for m in range(1, nadd + 1):
            p, c, s = out.struc(base + m)
            assert np.array_equal(c, c0) and np.array_equal(s, s0)
            dr = p - p0
            assert np.all(np.linalg.norm(dr, axis=1) <= 0.05 + 1e-15)
            assert E2[base + m] == pytest.approx(E[k] - np.sum(dr * f0), rel=0, abs=1e-12)
            assert not np.any(F2[oo[base + m]:oo[base + m + 1]])


def test_structures_without_forces_are_not_copied(oracle_mod):
    _, batch, _, E, F = _labelled(oracle_mod)
    o = batch.atom_off
    F = F.copy()
    F[o[0]:o[1]] = 0.0
    out, _, _, hasf = taylor_augment(batch, E, F, 2, 0.05, seed=2)
    assert out.nstruc == 1 + 1 + 2
    assert list(hasf) == [True, True, False, False]


def test_first_order_against_oracle(oracle_mod):
    _, batch, orc, E, F = _labelled(oracle_mod, nstruc=1)
    errs = []
    for h in (2e-3, 1e-3):
        out, E2, _, _ = taylor_augment(batch, E, F, 2, h, seed=3)
        Et, _, _ = orc.compute_batch(out, nthreads=4)
        lin = np.abs(E2[1:] - E[0])
        err = np.abs(np.asarray(Et)[1:] - E2[1:])
        assert np.all(err < 0.05 * lin)
        errs.append(err.max())
    # second-order remainder: a quarter when the displacement halves
    assert 0.15 < errs[1] / errs[0] < 0.35, errs


def test_directions_uniform_on_sphere():
    rng = np.random.default_rng(4)
    n = 20000
    b = configs.single(rng.uniform(0, 10, size=(n, 3)), np.eye(3) * 10.0, np.zeros(n, np.int32))
    F = np.ones((n, 3))
    out, _, _, _ = taylor_augment(b, np.zeros(1), F, 1, 1.0, dist="normal", seed=5)
    dr = out.struc(1)[0] - b.pos
    u = dr / np.linalg.norm(dr, axis=1)[:, None]
    assert np.all(np.abs(u.mean(axis=0)) < 0.03)
    assert np.allclose((u * u).mean(axis=0), 1.0 / 3.0, atol=0.02)


def test_unsupported_distribution_raises():
    b = configs.single([[0.0, 0.0, 0.0]], np.eye(3) * 5.0, [0])
    with pytest.raises(ValueError):
        taylor_augment(b, np.zeros(1), np.ones((1, 3)), 1, 0.1, dist="sech")
