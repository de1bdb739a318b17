"""Pins the oracle's neighbour list, symmetry functions, forces, virial,
tangent and training gradient (CPU only).

Checks (what each would catch):
* brute-force image enumeration = cell list, on wide, narrow, triclinic and
  unwrapped-position cells (a missed image or wrong shift);
* reverse slots form an involution with r(rev s) = -r(s) (a wrong sign or index);
* G invariant between the NaCl primitive cell and its 216-atom supercell, the
  fixtures of src/nnp/test_nnp_symm.cpp, and between a random triclinic cell and
  its 3x3x3 supercell (a dropped image, a wrong pair index);
* G of two- and three-atom clusters against the closed form (a transposed
  basis block, a missing cutoff factor, an r_jk term that must not be there);
* forces = -dE/dr by central differences (src/nnp/test_nnp_force.cpp pattern),
  sum F = 0, virial = -dE/dstrain by central differences;
* the reference's min()-gated reactions equal the exact adjoint for wide cells
  and differ for narrow ones;
* tangent = directional derivative of G; training gradient by central
  differences (h = 1e-4 as in the training workload definition).
"""
import json
import os

import numpy as np
import pytest

from ann_amd import configs
from ann_amd.model import ANG, CUT, NEURON, RAD, Model, init_network, pair_index

HERE = os.path.dirname(os.path.abspath(__file__))


def model3(rc=6.0, seed=0, cut=CUT["COS"], neuron="TANH", hidden=(5, 3)):
    """3-species potential in the shape of src/nnp/test_nnp_force.cpp:106-136
    with angular terms switched on."""
    rng = np.random.default_rng(seed)
    S = 3
    rad = (RAD["GAUSSIAN"], np.linspace(1.0, 5.0, 4), np.array([1.5, 1.0, 2.0, 0.8]))
    ang = (ANG["GAUSS"], np.array([2.0, 2.0, 4.0, 4.0]), np.array([1.0, 2.0, 1.0, 4.0]),
           np.array([1, -1, 1, -1]))
    radial = [[rad for _ in range(S)] for _ in range(S)]
    angular = [[ang for _ in range(S * (S + 1) // 2)] for _ in range(S)]
    nets = [init_network(4 * S + 4 * S * (S + 1) // 2, list(hidden), NEURON[neuron], rng, sigma_b=0.5)
            for _ in range(S)]
    return Model(["A", "B", "C"], np.ones(S), np.array([-1.0, -2.0, -3.0]), rc, cut, radial, angular, nets)


def structures():
    rng = np.random.default_rng(5)
    out = []
    out.append(configs.small_random(1, natoms=20, a0=20.0))               # wide cubic, unwrapped positions
    out.append(configs.small_random(2, natoms=12, a0=7.0))                # narrow: widths < 2 rc
    e = rng.uniform(-0.1, 0.1, (3, 3))
    cell = (np.eye(3) + 0.5 * (e + e.T)) @ np.array([[7.5, 1.0, 0.5], [0.0, 8.0, 1.2], [0.0, 0.0, 6.5]])
    pos = rng.uniform(-3, 12, (15, 3))
    out.append(configs.single(pos, cell, rng.integers(0, 3, 15)))          # triclinic, strained
    out.append(configs.single(rng.uniform(0, 4, (6, 3)), np.eye(3) * 4.0, rng.integers(0, 3, 6)))  # width < rc
    return out


@pytest.mark.parametrize("k", range(4))
def test_brute_force_equals_cell_list(oracle_mod, k):
    b = structures()[k]
    orc = oracle_mod.Oracle(model3())
    pos, cell, sp = b.struc(0)
    a = orc.nlist(pos, cell, sp, method=0)
    c = orc.nlist(pos, cell, sp, method=1)
    assert a.npairs > 0
    for f in ("off", "idx", "img", "rev", "minimg"):
        assert np.array_equal(getattr(a, f), getattr(c, f)), f
    assert np.array_equal(a.r, c.r)
    # canonical order: (j, n) strictly increasing per atom
    for i in range(len(pos)):
        keys = [tuple([a.idx[q]] + list(a.img[q])) for q in range(a.off[i], a.off[i + 1])]
        assert keys == sorted(keys) and len(set(keys)) == len(keys)
    # every listed vector is r_i - r_j - R n and inside the cutoff
    i_of = np.repeat(np.arange(len(pos)), np.diff(a.off))
    r = pos[i_of] - pos[a.idx] - a.img @ cell.T
    assert np.allclose(r, a.r, atol=1e-12) and np.all(a.d < 6.0) and np.all(a.d > 0)


@pytest.mark.parametrize("k", range(4))
def test_reverse_slots(oracle_mod, k):
    b = structures()[k]
    orc = oracle_mod.Oracle(model3())
    nl = orc.nlist(*b.struc(0))
    assert np.array_equal(nl.rev[nl.rev], np.arange(nl.npairs))
    assert np.array_equal(nl.r[nl.rev], -nl.r)
    i_of = np.repeat(np.arange(len(nl.off) - 1), np.diff(nl.off))
    assert np.array_equal(nl.idx[nl.rev], i_of)


def test_neighbor_count_brute_small_cell(oracle_mod):
    """A single atom in a small cubic cell sees exactly its lattice images
    within rc: count = #{n != 0 : |L n| < rc}."""
    L, rc = 2.5, 6.0
    m = model3(rc=rc)
    orc = oracle_mod.Oracle(m)
    nl = orc.nlist(np.zeros((1, 3)), np.eye(3) * L, [0])
    rng = np.arange(-4, 5)
    n = np.array([[a, b, c] for a in rng for b in rng for c in rng])
    want = np.sum((np.linalg.norm(n * L, axis=1) < rc) & (np.abs(n).sum(1) > 0))
    assert nl.npairs == want


def nacl_fixtures():
    d = json.load(open(os.path.join(HERE, "golden", "nacl.json")))
    out = []
    for key in ("small", "large"):
        f = d[key]
        a0 = f["a0"]
        cell = np.array(f["lv_rows"], float) * a0
        pos = np.array(f["pos_units_a0"], float) * a0
        out.append((pos, cell, np.array(f["species"], np.int32)))
    return out


def nacl_model():
    """The basis of src/nnp/test_nnp_symm.cpp:338-375: rc 8, COS cutoff,
    6 radial TANH (rs 1.06, eta = i) and 4 angular GAUSS (eta = zeta = i,
    lambda 1; i starts at 1 here because eta = 0 is singular)."""
    S = 2
    rad = (RAD["TANH"], np.full(6, 1.06), np.arange(6, dtype=float))
    ang = (ANG["GAUSS"], np.arange(1, 5, dtype=float), np.arange(1, 5, dtype=float), np.ones(4, int))
    radial = [[rad] * S for _ in range(S)]
    angular = [[ang] * 3 for _ in range(S)]
    rng = np.random.default_rng(0)
    nets = [init_network(6 * 2 + 4 * 3, [5, 3], NEURON["SWISH"], rng) for _ in range(S)]
    return Model(["Na", "Cl"], np.array([22.99, 35.45]), np.array([-1.68396789, -2.15869789]), 8.0,
                 CUT["COS"], radial, angular, nets)


def test_symm_supercell_invariance_nacl_fixture(oracle_mod):
    """NaCl 2-atom primitive fcc cell vs the 216-atom cubic supercell of
    src/nnp/test_nnp_symm.cpp (coordinates stored in single precision there,
    so the agreement is at the 1e-6 level)."""
    orc = oracle_mod.Oracle(nacl_model())
    (ps, cs, ss), (pl, cl, sl) = nacl_fixtures()
    Gs = orc.symm(ps, cs, ss)
    Gl = orc.symm(pl, cl, sl)
    scale = np.abs(Gs).max()
    for i in range(len(pl)):
        assert np.abs(Gl[i] - Gs[sl[i]]).max() < 2e-6 * scale


def test_symm_supercell_invariance_random_triclinic(oracle_mod):
    """Random 8-atom cell in an fcc-shaped lattice vs its 3x3x3 supercell
    (src/nnp/test_nnp_symm.cpp:387-500)."""
    rng = np.random.default_rng(11)
    a0 = 5.6199998856 * 2
    cell = np.array([[0.5, 0.5, 0.0], [0.5, 0.0, 0.5], [0.0, 0.5, 0.5]]) * a0
    pos = rng.uniform(-1, 1, (8, 3)) * a0
    sp = np.array([0, 1] * 4, np.int32)
    big_pos = np.concatenate([pos + cell @ np.array([i, j, k]) for i in range(3) for j in range(3) for k in range(3)])
    big_sp = np.tile(sp, 27)
    orc = oracle_mod.Oracle(nacl_model())
    Gs = orc.symm(pos, cell, sp)
    Gl = orc.symm(big_pos, cell * 3, big_sp)
    assert np.allclose(Gl, np.tile(Gs, (27, 1)), rtol=1e-10, atol=1e-12 * np.abs(Gs).max())


def test_symm_closed_form_clusters(oracle_mod):
    """Two and three atoms in a large box: G from the formulas by hand."""
    m = model3(rc=6.0)
    orc = oracle_mod.Oracle(m)
    box = np.eye(3) * 50.0
    fc = lambda r: 0.5 * (np.cos(np.pi * r / 6.0) + 1.0)
    # dimer A-B at 2.3 A
    G = orc.symm(np.array([[0, 0, 0], [2.3, 0, 0]]), box, [0, 1])
    _, rs, eta = m.radial[0][1]
    offr, offa = m.offsets(0)
    want = np.exp(-(eta * (2.3 - rs)) ** 2) * fc(2.3)
    assert np.allclose(G[0, offr[1]:offr[1] + 4], want, rtol=1e-9)
    assert np.all(G[0, offr[0]:offr[0] + 4] == 0) and np.all(G[0, offa[0]:] == 0)
    # triangle: centre A at origin, B and C at distances 1.7 and 2.9, angle 70 deg
    th = np.deg2rad(70)
    pos = np.array([[0, 0, 0], [1.7, 0, 0], [2.9 * np.cos(th), 2.9 * np.sin(th), 0]])
    G = orc.symm(pos, box, [0, 1, 2])
    _, etaA, zeta, lam = m.angular[0][pair_index(1, 2)]
    c = np.cos(th)
    wantA = np.abs(0.5 + 0.5 * lam * c) ** zeta * np.exp(-(1.7 ** 2 + 2.9 ** 2) / etaA ** 2) * fc(1.7) * fc(2.9)
    o = offa[pair_index(1, 2)]
    assert np.allclose(G[0, o:o + 4], wantA, rtol=1e-8)
    other = [offa[p] for p in range(6) if p != pair_index(1, 2)]
    assert all(np.all(G[0, q:q + 4] == 0) for q in other)


def energy(orc, pos, cell, sp):
    return orc.compute(pos, cell, sp)[0]


@pytest.mark.parametrize("k", [0, 1, 2, 3])
def test_forces_central_differences(oracle_mod, k):
    b = structures()[k]
    m = model3(seed=k)
    orc = oracle_mod.Oracle(m)
    pos, cell, sp = b.struc(0)
    E, _, F, V = orc.compute(pos, cell, sp)
    assert np.abs(F.sum(0)).max() < 1e-10 * max(1, np.abs(F).max())
    eps = 1e-6
    for i in range(min(len(pos), 6)):
        for x in range(3):
            p1, p2 = pos.copy(), pos.copy()
            p1[i, x] += eps
            p2[i, x] -= eps
            fd = -(energy(orc, p1, cell, sp) - energy(orc, p2, cell, sp)) / (2 * eps)
            assert abs(fd - F[i, x]) < 1e-6 * max(1.0, np.abs(F).max()), (i, x, fd, F[i, x])


@pytest.mark.parametrize("k", [0, 2])
def test_virial_strain_derivative(oracle_mod, k):
    b = structures()[k]
    orc = oracle_mod.Oracle(model3(seed=7))
    pos, cell, sp = b.struc(0)
    _, _, _, V = orc.compute(pos, cell, sp)
    eps = 1e-6
    for a in range(3):
        for c in range(3):
            D = np.zeros((3, 3))
            D[a, c] = eps
            Ep = energy(orc, pos @ (np.eye(3) + D).T, (np.eye(3) + D) @ cell, sp)
            Em = energy(orc, pos @ (np.eye(3) - D).T, (np.eye(3) - D) @ cell, sp)
            assert abs(-(Ep - Em) / (2 * eps) - V[a, c]) < 1e-6 * max(1, np.abs(V).max())


def test_reference_gating_matches_exact_for_wide_cells(oracle_mod):
    m = model3(seed=3)
    orc = oracle_mod.Oracle(m)
    wide = configs.small_random(9, natoms=20, a0=20.0)
    narrow = configs.small_random(10, natoms=10, a0=7.0)
    p, c, s = wide.struc(0)
    F0 = orc.compute(p, c, s, mode=0)[2]
    F1 = orc.compute(p, c, s, mode=1)[2]
    assert np.allclose(F0, F1, atol=1e-12)
    p, c, s = narrow.struc(0)
    F0 = orc.compute(p, c, s, mode=0)[2]
    F1 = orc.compute(p, c, s, mode=1)[2]
    assert np.abs(F0 - F1).max() > 1e-6      # dropped image reactions change the forces


def test_tangent_is_directional_derivative(oracle_mod):
    b = structures()[2]
    orc = oracle_mod.Oracle(model3(seed=4))
    pos, cell, sp = b.struc(0)
    u = np.random.default_rng(3).normal(size=pos.shape)
    v = orc.tangent(pos, cell, sp, u)
    h = 1e-6
    Gp = orc.symm(pos + h * u, cell, sp)
    Gm = orc.symm(pos - h * u, cell, sp)
    fd = (Gp - Gm) / (2 * h)
    assert np.allclose(v, fd, rtol=1e-5, atol=1e-8 * np.abs(v).max())


def test_training_gradient_central_differences(oracle_mod):
    """dL/dp of the energy + force loss by central differences, h = 1e-4."""
    m = model3(seed=8, neuron="TANH", hidden=(4,))
    b = configs.Batch.concat([configs.small_random(21, natoms=8, a0=9.0), configs.small_random(22, natoms=6, a0=8.0)])
    orc = oracle_mod.Oracle(m)
    rng = np.random.default_rng(4)
    Eref = rng.normal(size=2)
    Fref = rng.normal(size=(b.ntotal, 3)) * 0.1
    L, g, _, _, _ = orc.train_loss_grad(b, Eref, Fref, 1.0, 0.1)
    p0 = m.param_vector()
    h = 1e-4
    for k in rng.choice(len(p0), 15, replace=False):
        vals = []
        for s in (1, -1):
            p = p0.copy()
            p[k] += s * h
            m.set_param_vector(p)
            vals.append(oracle_mod.Oracle(m).train_loss_grad(b, Eref, Fref, 1.0, 0.1)[0])
        m.set_param_vector(p0)
        fd = (vals[0] - vals[1]) / (2 * h)
        assert abs(fd - g[k]) < 1e-6 * max(1.0, abs(g[k])), (k, fd, g[k])


@pytest.mark.parametrize("loss,norm", [(2, 1), (3, 1), (4, 1), (1, 0), (1, 2), (1, 3), (1, 4), (3, 2)])
def test_training_loss_variants_central_differences(oracle_mod, loss, norm):
    """The energy loss variants of the reference trainer (MAE 2, HUBER 3,
    ASINH 4; normalisations IDENTITY 0 .. LOG 4) with per-structure weights
    (an energy-only structure: force weight 0): dL/dp by central differences."""
    m = model3(seed=9, neuron="TANH", hidden=(4,))
    b = configs.Batch.concat([configs.small_random(23, natoms=8, a0=9.0), configs.small_random(24, natoms=6, a0=8.0),
                              configs.small_random(25, natoms=7, a0=8.5)])
    rng = np.random.default_rng(5)
    Eref = rng.normal(size=3) * 2.0
    Fref = rng.normal(size=(b.ntotal, 3)) * 0.1
    wEs, wFs = np.array([1.0, 0.5, 2.0]), np.array([1.0, 0.0, 1.5])
    kw = dict(loss=loss, delta=0.7, norm=norm, wEs=wEs, wFs=wFs)
    L, g, _, _, parts = oracle_mod.Oracle(m).train_loss_grad(b, Eref, Fref, 1.0, 0.1, **kw)
    assert np.isclose(L, parts.sum())
    p0 = m.param_vector()
    h = 1e-4
    for k in rng.choice(len(p0), 12, replace=False):
        vals = []
        for s in (1, -1):
            p = p0.copy()
            p[k] += s * h
            m.set_param_vector(p)
            vals.append(oracle_mod.Oracle(m).train_loss_grad(b, Eref, Fref, 1.0, 0.1, **kw)[0])
        m.set_param_vector(p0)
        fd = (vals[0] - vals[1]) / (2 * h)
        assert abs(fd - g[k]) < 1e-6 * max(1.0, abs(g[k])), (k, fd, g[k])


def test_training_loss_limits(oracle_mod):
    """HUBER and ASINH tend to MSE for a large width; a structure with weights
    0 contributes nothing; MAE of one structure is |dE| norm."""
    m = model3(seed=9, neuron="TANH", hidden=(4,))
    b = configs.Batch.concat([configs.small_random(23, natoms=8, a0=9.0), configs.small_random(24, natoms=6, a0=8.0)])
    rng = np.random.default_rng(6)
    Eref = rng.normal(size=2)
    Fref = rng.normal(size=(b.ntotal, 3)) * 0.1
    orc = oracle_mod.Oracle(m)
    Lm, gm, E, _, pm = orc.train_loss_grad(b, Eref, Fref, 1.0, 0.1)
    for loss in (3, 4):
        L, g, _, _, p = orc.train_loss_grad(b, Eref, Fref, 1.0, 0.1, loss=loss, delta=1e4)
        assert np.isclose(p[0], pm[0], rtol=1e-6) and np.allclose(g, gm, rtol=1e-6, atol=1e-12)
    L0, g0, _, _, _ = orc.train_loss_grad(b, Eref, Fref, 1.0, 0.1, wEs=[1.0, 0.0], wFs=[1.0, 0.0])
    L1, g1, _, _, _ = orc.train_loss_grad(b.subset([0]), Eref[:1], Fref[:8], 1.0, 0.1)
    assert np.isclose(L0, L1 / 2) and np.allclose(g0, g1 / 2, rtol=1e-12, atol=1e-15)
    _, _, _, _, pa = orc.train_loss_grad(b, Eref, Fref, 1.0, 0.1, loss=2, norm=0)
    assert np.isclose(pa[0], np.abs(E - Eref).sum() / 2)
