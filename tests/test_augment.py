"""Taylor augmentation (src/nnp/nnptefr2.cpp:2383-2482 restated in
ann_amd/augment.py), pinned on the CPU:

* layout: each original followed by its nadd copies, copies without forces,
  structures without forces not copied (a wrong insertion order or count);
* E' = E - sum dr . F exactly, recomputed from the copies' positions (a sign or
  a dropped atom in the Taylor term);
* against the FP64 oracle the estimate is first-order: E(r + dr) - E' shrinks
  fourfold when the displacement halves, and is far below the linear term
  (a displacement applied with the wrong sign or to the wrong atom breaks this);
* directions uniform on the sphere (a biased theta or phi draw);
* the reference stream (ann_amd/refrng.py): std::mt19937 against the C++
  standard's known answer (the 10000th output of a default-constructed engine
  is 4123659995, [rand.predef]) and, bit for bit, against the reference's own
  src/rand/rng.cpp and the standard library compiled into oracle/_ref (raw
  engine, uniform_real_distribution, all six distributions, the trainer's
  seed chain and per-atom draw order of src/nnp/nnptefr2.cpp:2160-2168,
  2447-2458); the positions and energies of the copies rebuilt from those
  draws; one stream continued over two sets equal to one pass over both;
  sample moments of each distribution against its closed form (a wrong scale
  or a dropped factor).
"""
import numpy as np
import pytest

from ann_amd import configs
from ann_amd.augment import TrainerDraws, taylor_augment
from ann_amd.refrng import MT19937, Distribution, trainer_generator, uniform_real


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
    with pytest.raises(ValueError):
        taylor_augment(b, np.zeros(1), np.ones((1, 3)), 1, 0.1, dist="cauchy", rng="reference", seed=3)
    with pytest.raises(ValueError):      # the reference seeds from the clock for seed 0
        taylor_augment(b, np.zeros(1), np.ones((1, 3)), 1, 0.1, rng="reference", seed=0)
    with pytest.raises(ValueError):
        taylor_augment(b, np.zeros(1), np.ones((1, 3)), 1, 0.1, rng="mt")


# ---- the reference's random stream ----

DISTS = Distribution.NAMES      # rng::dist numbering of the wrapper: 0..5


def test_mt19937_known_answer():
    g = MT19937()
    for _ in range(9999):
        g()
    assert g() == 4123659995
    assert MT19937(5489)() == 3499211612      # first output of the default seed


def _ref_rng(oracle_mod):
    L = oracle_mod.ref_rng_lib()
    if L is None:
        pytest.skip("oracle/_ref not built (reference tree absent)")
    return L


@pytest.mark.parametrize("seed", [1, 5489, 2**32 - 1])
def test_engine_and_uniform_match_reference_bitwise(oracle_mod, seed):
    L = _ref_rng(oracle_mod)
    n = 2000                                   # crosses three twists
    raw = np.zeros(n, np.uint32)
    L.ref_mt_raw(seed, n, raw.ctypes.data)
    g = MT19937(seed)
    assert np.array_equal(raw, np.array([g() for _ in range(n)], np.uint32))
    for a, b in ((-1.0, 1.0), (0.0, 0.05)):
        ref = np.zeros(n)
        L.ref_uniform(seed, a, b, n, ref.ctypes.data)
        g = MT19937(seed)
        assert np.array_equal(ref, np.array([uniform_real(g, a, b) for _ in range(n)]))


@pytest.mark.parametrize("k", range(6))
def test_distributions_match_reference_bitwise(oracle_mod, k):
    L = _ref_rng(oracle_mod)
    n = 3001                                   # odd: Normal ends on a fresh pair
    for s, seed in ((0.05, 11), (1.7, 4242)):
        ref = np.zeros(n)
        L.ref_rng_draw(k, s, seed, n, ref.ctypes.data)
        g, d = MT19937(seed), Distribution(DISTS[k], s)
        assert np.array_equal(ref, np.array([d.rand(g) for _ in range(n)])), DISTS[k]


@pytest.mark.parametrize("k", range(6))
def test_trainer_chain_and_draw_order_match_reference(oracle_mod, k):
    L = _ref_rng(oracle_mod)
    natoms = 301                               # odd: Normal's cached partner crosses atoms
    ref = np.zeros(3 * natoms)
    L.ref_taylor_draws(k, 0.03, 987654321, natoms, ref.ctypes.data)
    t = TrainerDraws(987654321, DISTS[k], 0.03)
    assert np.array_equal(ref.reshape(-1, 3), np.array([t.atom() for _ in range(natoms)]))


def test_reference_stream_copies_rebuilt_from_reference_draws(oracle_mod):
    L = _ref_rng(oracle_mod)
    _, batch, _, E, F = _labelled(oracle_mod)
    nadd, seed, rdelta = 2, 31337, 0.04
    out, E2, F2, hasf = taylor_augment(batch, E, F, nadd, rdelta, dist="normal", seed=seed, rng="reference")
    o = batch.atom_off
    nd = nadd * int(o[-1])
    draws = np.zeros(3 * nd)
    L.ref_taylor_draws(2, rdelta, seed, nd, draws.ctypes.data)
    draws = draws.reshape(-1, 3)
    q = 0
    for k in range(batch.nstruc):
        p0, f0 = batch.struc(k)[0], F[o[k]:o[k + 1]]
        for m in range(1, nadd + 1):
            vr, ct, u = draws[q:q + len(p0)].T
            q += len(p0)
            st = np.sqrt(1.0 - ct * ct)
            dr = vr[:, None] * np.stack([st * np.cos(u * np.pi), st * np.sin(u * np.pi), ct], axis=1)
            p = out.struc(k * (1 + nadd) + m)[0]
            assert np.allclose(p, p0 + dr, rtol=0, atol=1e-15)
            assert E2[k * (1 + nadd) + m] == pytest.approx(E[k] - np.sum(dr * f0), rel=0, abs=1e-12)
    assert list(hasf) == ([True] + [False] * nadd) * batch.nstruc


def test_reference_stream_continues_over_sets(oracle_mod):
    _, batch, _, E, F = _labelled(oracle_mod, nstruc=3)
    o = batch.atom_off
    sets = [batch.subset(range(a, b)) for a, b in ((0, 1), (1, 3))]
    labels = [(E[a:b], F[o[a]:o[b]]) for a, b in ((0, 1), (1, 3))]
    whole = taylor_augment(batch, E, F, 2, 0.05, dist="normal", seed=77, rng="reference")
    t = TrainerDraws(77, "normal", 0.05)
    parts = [taylor_augment(b, e, f, 2, 0.05, rng="reference", draws=t) for b, (e, f) in zip(sets, labels)]
    assert np.array_equal(whole[0].pos, np.concatenate([p[0].pos for p in parts]))
    assert np.array_equal(whole[1], np.concatenate([p[1] for p in parts]))


def test_trainer_seed_chain():
    # rngen = mt19937(seed_global); seed_local = rngen(); rngen = mt19937(seed_local)
    g0 = MT19937(12345)
    g = MT19937(g0())
    h = trainer_generator(12345)
    assert [g() for _ in range(700)] == [h() for _ in range(700)]


@pytest.mark.parametrize("name,var", [
    ("laplace", 2.0), ("normal", 1.0), ("logistic", np.pi ** 2 / 3.0), ("sech", 1.0)])
def test_distribution_moments(name, var):
    g, d = MT19937(2024), Distribution(name, 0.5)
    x = np.array([d.rand(g) for _ in range(60000)])
    assert abs(x.mean()) < 0.02
    assert x.var() == pytest.approx(var * 0.25, rel=0.04)


def test_uniform_and_cosine_ranges():
    g = MT19937(7)
    u = np.array([Distribution("uniform", 0.3).rand(g) for _ in range(20000)])
    assert u.min() >= 0.0 and u.max() <= 0.3 and u.mean() == pytest.approx(0.15, abs=0.005)
    # raised cosine on [-s, s] (an approximate inverse CDF): symmetric, variance s^2 (1/3 - 2/pi^2)
    d = Distribution("cosine", 1.0)
    c = np.array([d.rand(g) for _ in range(40000)])
    assert abs(c.mean()) < 0.01
    assert c.var() == pytest.approx(1.0 / 3.0 - 2.0 / np.pi ** 2, rel=0.05)
